"""Write tests/golden/mt19937_ref.json from the reference's own MT19937 (mtprng.f90),
compiled by `make -C oracle ref` into oracle/_ref/mt_ref (run in the build container,
where the ESpinS sources are present).  Fixture = seeds, counts, outputs."""
import json
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "..", "..", "oracle", "_ref", "mt_ref")


def main():
    out = {"_source": "mtprng_rand64 / mtprng_rand_real2 of mtprng.f90 (ESpinS), compiled from "
                      "its own source by oracle/Makefile `ref` with oracle/mt_ref_harness.f90"}
    for seed, n in ((5489, 1300), (12345, 700), (2024, 700), (1, 700)):
        txt = subprocess.run([EXE, str(seed), str(n)], capture_output=True, text=True,
                             check=True).stdout.split()
        out[str(seed)] = {"u32": [int(x) for x in txt[:n]],
                          "real2": [float(x) for x in txt[n:2 * n]]}
    with open(os.path.join(HERE, "mt19937_ref.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
