"""Writes tests/golden/cereal/{nvt,gcmc}_restart.{chk,json} with
oracle/_ref/ref_cereal_chk, the program compiled (by `make -C oracle ref`)
against the reference's vendored cereal (lib/cereal) and its MolSetup
serializers (src/MolSetup.h): the archive bytes cereal writes for a
checkpoint of one or two boxes, and the values it wrote.  Run in a container
where the reference tree exists; the files it writes are the committed
fixtures."""
import json
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
OUT = HERE / "cereal"


def main():
    subprocess.run(["make", "-C", str(ROOT / "oracle"), "ref"], check=True)
    OUT.mkdir(exist_ok=True)
    for name, nb in (("nvt", 1), ("gcmc", 2)):
        chk = OUT / f"{name}_restart.chk"
        out = subprocess.run([str(ROOT / "oracle" / "_ref" / "ref_cereal_chk"), str(nb),
                              str(chk)], check=True, capture_output=True, text=True).stdout
        data = json.loads(out)
        data["archive"] = chk.name
        (OUT / f"{name}_restart.json").write_text(json.dumps(data, indent=1) + "\n")
        print("wrote", chk, "and its values")


if __name__ == "__main__":
    main()
