"""Write tests/golden/ran3_golden.json from the reference's own RNG code.

Runs oracle/_ref/ref_psnr, which `make -C oracle ref` compiles from the
reference's src/PSNR/PSNR.cpp and src/PSNR/ran3.c (see oracle/Makefile).
The JSON holds, per seed: the 55 state words and (inext, inextp) that
PSNR::InitRand3 / GetRan3Array / GetInextInextp hand to the kernel, the next
200 integer draws of PSNR::GetRan3, and the first 201 draws of ran3().
"""
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SEEDS = [1, 2, 7, 12345, 32767, 1000000, 161803397]


def main():
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_psnr")
    if not os.path.exists(exe):
        sys.exit("build it first: make -C oracle ref")
    out = subprocess.run([exe] + [str(s) for s in SEEDS], check=True,
                         capture_output=True, text=True).stdout
    data = json.loads(out)
    with open(os.path.join(HERE, "ran3_golden.json"), "w") as f:
        json.dump({"source": "src/PSNR/PSNR.cpp + src/PSNR/ran3.c (reference), via oracle/_ref/ref_psnr",
                   "cases": data}, f, separators=(",", ":"))
    print("wrote", len(data), "cases")


if __name__ == "__main__":
    main()
