"""Writes the dump-format fixtures from the reference's shipped data files (run in a container
that has the reference at /root/reference; the outputs are committed):

  asph_t000060s.bin, asph_t000300s.bin  - two of dumps_old/ (128x128x64 and 160x160x80 runs)
  dumps_new_t300.bin                    - the 128x128x64 dump at t = 300 s
  viz_frames.json                       - the frames of sph_data.json (the visualiser's
                                          conversion of dumps_old/) for those two dumps
"""
import json
import os
import shutil

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

FILES = ["asph_t000060s.bin", "asph_t000300s.bin"]


def main():
    for f in FILES:
        shutil.copyfile(os.path.join(REF, "dumps_old", f), os.path.join(HERE, f))
    shutil.copyfile(os.path.join(REF, "dumps_new_t300.bin"), os.path.join(HERE, "dumps_new_t300.bin"))
    data = json.load(open(os.path.join(REF, "sph_data.json")))
    all_files = sorted(os.listdir(os.path.join(REF, "dumps_old")))
    out = {"top": {k: data[k] for k in ("Lx", "Ly", "Lz", "cx", "cy", "cz", "dx", "dy", "dz", "theta_ref")},
           "frames": {}}
    for f in FILES:
        k = all_files.index(f)
        out["frames"][f] = {"t": data["times"][k], "theta": data["theta"][k], "qv": data["qv"][k],
                            "ql": data["ql"][k], "w": data["w"][k]}
    with open(os.path.join(HERE, "viz_frames.json"), "w") as fh:
        json.dump(out, fh, separators=(",", ":"))


if __name__ == "__main__":
    main()
