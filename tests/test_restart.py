"""Restart records against files written by the reference's write_restart_files (host code, CPU)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from freecappuccino_amd import fcap, restart


@pytest.fixture(scope="module")
def ref():
    return np.load(os.path.join(GOLDEN, "restart_ref.npz"), allow_pickle=False)


def _fields(z, tag):
    return {k: z[f"{tag}_{k}"] for k in restart.FIELDS}


@pytest.mark.parametrize("tag", ["plain", "mflux"])
def test_writer_reproduces_reference_file(ref, tag, tmp_path):
    nc, nif, ntot, cm, itime = (int(x) for x in ref[f"{tag}_meta"])
    time, g = ref[f"{tag}_scal"]
    path = str(tmp_path / "restart")
    restart.write(path, _fields(ref, tag), itime=itime, time=float(time), gradpcmf=float(g) if cm else None)
    assert open(path, "rb").read() == ref[f"{tag}_file"].tobytes()


@pytest.mark.parametrize("tag", ["plain", "mflux"])
def test_reader_reads_reference_file(ref, tag, tmp_path):
    nc, nif, ntot, cm, itime = (int(x) for x in ref[f"{tag}_meta"])
    path = str(tmp_path / "restart_ref")
    with open(path, "wb") as f:
        f.write(ref[f"{tag}_file"].tobytes())
    it, t, g, fields = restart.read(path, const_mflux=bool(cm))
    assert it == itime and t == ref[f"{tag}_scal"][0]
    if cm:
        assert g == ref[f"{tag}_scal"][1]
    assert list(fields) == list(restart.FIELDS)
    for k in restart.FIELDS:
        np.testing.assert_array_equal(fields[k], ref[f"{tag}_{k}"], err_msg=k)
    assert fields["flmass"].size == nif and fields["uu"].size == nc and fields["u"].size == ntot


def test_reader_rejects_damaged_files(ref, tmp_path):
    data = ref["plain_file"].tobytes()
    cut = str(tmp_path / "cut")
    with open(cut, "wb") as f:
        f.write(data[:-5])                       # last record's trailer cut short
    with pytest.raises(fcap.FcError):
        restart.read(cut)
    bad = bytearray(data)
    bad[-4] ^= 0xff                              # trailer differs from the header
    with open(str(tmp_path / "bad"), "wb") as f:
        f.write(bytes(bad))
    with pytest.raises(fcap.FcError):
        restart.read(str(tmp_path / "bad"))
    plain = str(tmp_path / "plain")
    with open(plain, "wb") as f:
        f.write(data)
    with pytest.raises(fcap.FcError):            # a constant-mass-flow reader on a plain file
        restart.read(plain, const_mflux=True)
    with pytest.raises(fcap.FcError):
        restart.read(str(tmp_path / "missing"))


def test_round_trip_large_records(tmp_path):
    rng = np.random.default_rng(0)
    f = {k: rng.standard_normal(1000 + 17 * i) for i, k in enumerate(restart.FIELDS)}
    path = str(tmp_path / "rt")
    restart.write(path, f, itime=3, time=0.25, gradpcmf=-1.5)
    it, t, g, back = restart.read(path, const_mflux=True)
    assert (it, t, g) == (3, 0.25, -1.5)
    for k in restart.FIELDS:
        np.testing.assert_array_equal(back[k], f[k])
