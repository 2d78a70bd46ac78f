"""Copies the reference's only fixtures, its JSON config files, into
tests/golden/reference_configs/ (same relative paths):

    ConfigTemplates/**/*.json            solver, object-component and generator templates
    Examples/<Solver>/Scene-1/Configuration.json   the four example scenes

They are data (solver parameters and object descriptions), used by the tests
to pin the config templates and to run the reference's own scenes.  Run in a
checkout that has the reference at REFERENCE_ROOT:

    python tests/golden/copy_reference_configs.py [REFERENCE_ROOT]
"""
import shutil
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "reference_configs"


def main(root: str = "/root/reference"):
    src = Path(root)
    files = sorted(src.glob("ConfigTemplates/**/*.json")) + sorted(src.glob("Examples/*/Scene-1/Configuration.json"))
    if not files:
        raise SystemExit(f"no reference config files under {src}")
    for f in files:
        dst = OUT / f.relative_to(src)
        dst.parent.mkdir(parents=True, exist_ok=True)
        shutil.copyfile(f, dst)
        print(dst.relative_to(HERE))


if __name__ == "__main__":
    main(*sys.argv[1:])
