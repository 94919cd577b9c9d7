"""Copies the reference's twelve preset documents (editor/presets/*.json, format fire-pro
version 11) into tests/golden/presets/ as input fixtures, after checking that each one
passes validate_document. Run where the reference is checked out:

    python tests/golden/make_presets.py /path/to/threejs-fire-pro
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "threejs-fire-pro_amd"))

from firepro.document import validate_document  # noqa: E402


def main(reference: str) -> None:
    src = os.path.join(reference, "editor", "presets")
    out = os.path.join(HERE, "presets")
    os.makedirs(out, exist_ok=True)
    for name in sorted(os.listdir(src)):
        if not name.endswith(".json"):
            continue
        with open(os.path.join(src, name)) as f:
            doc = json.load(f)
        validate_document(doc)
        with open(os.path.join(out, name), "w") as f:
            json.dump(doc, f, indent=1, sort_keys=False)
            f.write("\n")
        print(name)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
