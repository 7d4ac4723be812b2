"""Small datasets for the batched and served tests: two smooth seeded
volumes over the unit cube (compositing, iso 0.5) and two blob meshes."""
import numpy as np

from galaxy_amd import scenes, workloads

f32 = np.float32


def smooth_volume(n, seed):
    """n^3 voxels of a smooth seeded field in [0, 1] over the unit cube."""
    rng = np.random.default_rng(seed)
    x = np.linspace(0, 1, n, dtype=np.float64)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    f = np.zeros_like(X)
    for k in range(4):
        a = rng.normal(size=3)
        f += np.sin(2.5 * (k + 1) * (a[0] * X + a[1] * Y + a[2] * Z) + rng.uniform(0, 6.3)) / (k + 1)
    f = (f - f.min()) / (f.max() - f.min())
    delta = f32(1.0 / (n - 1))
    return scenes.VolumeDesc(f.astype(f32), (f32(0),) * 3, (delta,) * 3, sampling_rate=0.5,
                             local_box=workloads.UNIT_BOX, global_box=workloads.UNIT_BOX)


def dataset_descs():
    d = np.linspace(0.0, 1.0, 8)
    rgb, op = workloads.VOLUME_RENDER_TF
    out = []
    for k, n in enumerate((40, 56)):
        vol = smooth_volume(n, 10 + k)
        vv = scenes.VolumeVisDesc(vol, colormap=np.c_[d, rgb].astype(f32), opacitymap=np.c_[d, 4 * op].astype(f32),
                                  isovalues=np.array([0.5], f32), volume_render=True)
        out.append(scenes.SceneDesc(volume_vis=[vv], local_box=workloads.UNIT_BOX, global_box=workloads.UNIT_BOX))
    out.append(workloads.blob_scene(3000, 7))
    out.append(workloads.blob_scene(20000, 8))
    return out
