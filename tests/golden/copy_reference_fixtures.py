"""Copies the reference's Cornell reference-image fixtures into this directory: the scene
description scenes/S02_CornellDiffuse_64x64.json and the image it was rendered to,
scenes/references/S02_CornellDiffuse_64x64_scalar_ref_4096spp.exr (4096 spp by the reference's
scalar backend), each checked against the SHA-256 the reference's
scenes/references/S02_CornellDiffuse.references.json records. Both are data files the reference's
own tests read (Tests/Renderer/CornellDiffuseReferenceTest.cpp).

    python tests/golden/copy_reference_fixtures.py [/root/reference]
"""
import hashlib
import json
import shutil
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent


def main(reference):
    reference = Path(reference)
    manifest = json.loads((reference / "scenes/references/S02_CornellDiffuse.references.json").read_text())
    entry = next(r for r in manifest["references"] if r["variant"] == "64x64")
    for key, digest in (("scene", entry["scene_sha256"]), ("image", entry["image_sha256"])):
        source = reference / entry[key]
        got = hashlib.sha256(source.read_bytes()).hexdigest()
        if got != digest:
            raise SystemExit(f"{source}: sha256 {got}, the manifest records {digest}")
        shutil.copyfile(source, HERE / source.name)
        print(f"copied {source.name} ({source.stat().st_size} bytes, sha256 {got})")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
