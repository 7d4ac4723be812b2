"""Writes the reference-held fixtures this test suite checks against.

Run once in a container that has the TACC/Galaxy tree at /root/reference:
  * golds/<test>_<k>.png  -- tests/golds/*.png (Galaxy's image gold tests,
                             tests/image-gold-tests.sh:40-174), except
                             data-driven (its inputs need VTK)
  * states/<test>.state   -- tests/*.state, the JSON inputs of those images
  * uv_table.json         -- the 2 x 256 values of randomU / randomV
                             (src/renderer/UV.ih:21-58) as numbers
"""
import json
import os
import re
import shutil
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
TESTS = ["absolute", "absolute-shadow", "camera", "camera-shadow", "infinite", "infinite-shadow",
         "nineBalls", "oneBall", "xyz"]


def main():
    os.makedirs(os.path.join(HERE, "golds"), exist_ok=True)
    os.makedirs(os.path.join(HERE, "states"), exist_ok=True)
    for t in TESTS:
        shutil.copy(os.path.join(REF, "tests", t + ".state"), os.path.join(HERE, "states", t + ".state"))
        k = 0
        while os.path.exists(os.path.join(REF, "tests", "golds", "%s_%05d.png" % (t, k))):
            shutil.copy(os.path.join(REF, "tests", "golds", "%s_%05d.png" % (t, k)),
                        os.path.join(HERE, "golds", "%s_%05d.png" % (t, k)))
            k += 1
    txt = open(os.path.join(REF, "src", "renderer", "UV.ih")).read()
    out = {}
    for name in ("randomU", "randomV"):
        m = re.search(name + r"\[256\]\s*=\s*\{([^}]*)\}", txt)
        out[name] = [float(s) for s in m.group(1).split(",") if s.strip()]
    json.dump(out, open(os.path.join(HERE, "uv_table.json"), "w"))


if __name__ == "__main__":
    main()
