"""Copy the polyMesh data of the reference's example cases into small fixtures.

    python tests/golden/make_polymesh_fixtures.py

The reference ships two meshes as data (examples/cavity/cavity-setup.tar.gz,
examples/pitzDaily/pitzDaily-setup.tar.gz).  The five files the mesh reader
reads (points, faces, owner, neighbour, boundary) are taken out unchanged:
cavity as plain files under polymesh_cavity/, pitzDaily as a gzipped tar
(polymesh_pitzDaily.tar.gz).  Only data is copied, nothing is run.
"""
import io
import os
import tarfile

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/examples"
FILES = ("points", "faces", "owner", "neighbour", "boundary")


def members(tar):
    out = {}
    for m in tar.getmembers():
        parts = m.name.strip("./").split("/")
        if len(parts) >= 2 and parts[-2] == "polyMesh" and parts[-1] in FILES and m.isfile():
            out[parts[-1]] = m
    missing = set(FILES) - set(out)
    if missing:
        raise RuntimeError(f"polyMesh files missing: {sorted(missing)}")
    return out


def main():
    with tarfile.open(os.path.join(REF, "cavity", "cavity-setup.tar.gz")) as t:
        d = os.path.join(HERE, "polymesh_cavity")
        os.makedirs(d, exist_ok=True)
        for name, m in members(t).items():
            with open(os.path.join(d, name), "wb") as f:
                f.write(t.extractfile(m).read())
    with tarfile.open(os.path.join(REF, "pitzDaily", "pitzDaily-setup.tar.gz")) as t, \
            tarfile.open(os.path.join(HERE, "polymesh_pitzDaily.tar.gz"), "w:gz") as out:
        for name, m in sorted(members(t).items()):
            data = t.extractfile(m).read()
            info = tarfile.TarInfo(f"polyMesh/{name}")
            info.size = len(data)
            out.addfile(info, io.BytesIO(data))


if __name__ == "__main__":
    main()
