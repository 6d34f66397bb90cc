"""Writes tests/golden/ref_rng_kat.json from oracle/_ref/ref_rng_kat, the
driver compiled (by `make -C oracle ref`) against the reference's own RNG
sources: src/Random123Wrapper.cpp, lib/Random123/philox.h + uniform.hpp,
src/MersenneTwister.h and src/TransformMatrix.h.  Run in a container where
the reference tree exists; the JSON it writes is the committed fixture."""
import json
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent


def main():
    subprocess.run(["make", "-C", str(ROOT / "oracle"), "ref"], check=True)
    out = subprocess.run([str(ROOT / "oracle" / "_ref" / "ref_rng_kat")], check=True,
                         capture_output=True, text=True).stdout
    # MTRand::seed prints the seed to stdout (src/MersenneTwister.h:226-231)
    out = "\n".join(l for l in out.splitlines() if not l.startswith("Random number seed:"))
    data = json.loads(out)
    (HERE / "ref_rng_kat.json").write_text(json.dumps(data, indent=1) + "\n")
    print("wrote", HERE / "ref_rng_kat.json")


if __name__ == "__main__":
    main()
