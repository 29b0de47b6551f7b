"""Writes the golden fixtures of tests/golden from the reference's own data files.

Run in a container where the reference is mounted (it is not needed afterwards):
    python tests/golden/make_golden.py /root/reference

  randlib_seed1.json       the first 20 Park-Miller states from seed 1, the
                           known-answer table in Source_Code_4D_Fast/randlib.c:69-91
  interlaced_clip.json     the 399 views (angles, times) deleted by
                           clip_list_of_views in the captured launcher log
                           Source_Code_4D_Fast/info.txt:1-175 (K=32, N_theta=1536)
  view_info_reference.txt  the view_info.txt data file shipped with the engine
                           (identical in Source_Code_4D_Fast, _4D, _3D, _APS2014)
"""
import json
import os
import re
import shutil
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
F = os.path.join(ref, "Source_Code_4D_Fast")
out = os.path.dirname(os.path.abspath(__file__))

# randlib known-answer table (a comment block of 20 integers)
src = open(os.path.join(F, "randlib.c")).read()
block = src[src.index("TABLE:"):]
block = block[:block.index("*/")]
nums = [int(t) for t in re.findall(r"^\s*(\d+)\s*$", block, re.M)]
assert len(nums) == 20, nums
json.dump({"seed": 1, "states": nums, "source": "Source_Code_4D_Fast/randlib.c:69-91"},
          open(os.path.join(out, "randlib_seed1.json"), "w"), indent=1)

# deleted views and their times from the launcher log
log = open(os.path.join(F, "info.txt")).read()
def grab(tag):
    i = log.index(tag)
    j = log.index("[", i)
    k = log.index("]", j)
    return [float(t) for t in log[j + 1:k].split()]
views = grab("Deleted views are")
times = grab("Deleted views' times are")
m = re.search(r"K_(\d+)_N_theta_(\d+)_N_p_(\d+)", log)
assert len(views) == 399 and len(times) == 399
json.dump({"K": int(m.group(1)), "N_theta": int(m.group(2)), "N_p_used": int(m.group(3)),
           "deleted_views": views, "deleted_times": times,
           "source": "Source_Code_4D_Fast/info.txt:1-175"},
          open(os.path.join(out, "interlaced_clip.json"), "w"))

shutil.copyfile(os.path.join(F, "view_info.txt"), os.path.join(out, "view_info_reference.txt"))
print("wrote", sorted(os.listdir(out)))
