"""The FBH5 writer plugin of the rawspec program (librawspec_fbh5.so) on the
CPU, read back with HDF5's h5dump: the file layout the reference's
fbh5_open.c / fbh5_write.c / fbh5_close.c / fbh5_util.c produce (file
attributes, data set [time][nifs][nchans] float32 growing by whole
integrations, chunks of (spectra per dump, 1, fine channels per coarse
channel), header attributes with refbeam only when set, dimension labels)."""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import rawspec_amd as R  # noqa: E402
from tests import fbh5_read as F  # noqa: E402

PLUGIN = os.path.join(os.path.dirname(R.lib_path()), "bin", "librawspec_fbh5.so")

pytestmark = pytest.mark.skipif(not (os.path.exists(PLUGIN) and F.available()),
                                reason="FBH5 plugin or h5dump not available")


@pytest.fixture(scope="module")
def P():
    L = C.CDLL(PLUGIN)
    L.fbh5w_open.restype = C.c_void_p
    L.fbh5w_open.argtypes = [C.POINTER(R.FbHdr), C.c_uint, C.c_char_p, C.c_char_p, C.c_char_p,
                             C.c_char_p]
    L.fbh5w_write.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
    L.fbh5w_close.argtypes = [C.c_void_p]
    return L


def _hdr(**kw):
    h = R.FbHdr()
    h.machine_id, h.telescope_id, h.data_type, h.nbeams, h.ibeam = 20, 6, 1, 1, 3
    h.nbits, h.nifs, h.nchans, h.nfpc = 32, 4, 2048, 512
    h.fch1, h.foff, h.tstart, h.tsamp = 1501.25, -0.0234375, 59000.5, 1.073741824
    h.src_raj, h.src_dej, h.az_start, h.za_start = 17.1678, 12.4117, 0.0, 0.0
    h.source_name, h.rawdatafile, h.refbeam = b"VOYAGER1", b"guppi.0000.raw", -1
    for k, v in kw.items():
        setattr(h, k, v)
    return h


@pytest.mark.parametrize("nifs,nds,dumps", [(4, 3, 3), (1, 1, 5), (1, 16, 2)])
def test_layout_and_data(P, tmp_path, nifs, nds, dumps):
    h = _hdr(nifs=nifs)
    path = str(tmp_path / "x.rawspec.0000.h5")
    c = P.fbh5w_open(C.byref(h), nds, path.encode(), b"r1", b"l2", b"f3")
    assert c
    rng = np.random.default_rng(nifs * 100 + nds)
    chunks = [rng.standard_normal(nds * nifs * h.nchans).astype(np.float32) for _ in range(dumps)]
    for d in chunks:
        assert P.fbh5w_write(c, d.ctypes.data, d.nbytes) == 0
    assert P.fbh5w_close(c) == 0
    fa, da, info, data = F.read(path)
    assert fa["CLASS"] == "FILTERBANK" and fa["VERSION"] == "2.0"
    assert (fa["VERSION_RAWSPEC"], fa["VERSION_LIBRAWSPEC"], fa["VERSION_CUFFT"]) == ("r1", "l2", "f3")
    assert fa["BITSHUFFLE"] in ("ENABLED", "DISABLED") and fa["VERSION_HDF5"].count(".") == 2
    assert info["shape"] == (nds * dumps, nifs, h.nchans)
    assert info["chunk"] == (nds, 1, h.nfpc)
    assert info["dtype"] == "H5T_IEEE_F32LE"
    assert np.array_equal(data.reshape(-1), np.concatenate(chunks))
    for k in ("machine_id", "telescope_id", "data_type", "nchans", "nfpc", "nbeams", "ibeam",
              "nbits", "nifs"):
        assert da[k] == getattr(h, k), k
    for k in ("src_raj", "src_dej", "az_start", "za_start", "fch1", "foff", "tstart", "tsamp"):
        assert da[k] == getattr(h, k), k
    assert da["source_name"] == "VOYAGER1" and da["rawdatafile"] == "guppi.0000.raw"
    assert "refbeam" not in da
    assert da["DIMENSION_LABELS"] == ["time", "feed_id", "frequency"]


def test_refbeam_written_when_set(P, tmp_path):
    h = _hdr(refbeam=5)
    path = str(tmp_path / "r.h5")
    c = P.fbh5w_open(C.byref(h), 1, path.encode(), b"a", b"b", b"c")
    d = np.ones(4 * h.nchans, dtype=np.float32)
    assert P.fbh5w_write(c, d.ctypes.data, d.nbytes) == 0
    assert P.fbh5w_close(c) == 0
    _, da, info, data = F.read(path)
    assert da["refbeam"] == 5 and info["shape"] == (1, 4, h.nchans)


def test_bad_header_is_refused(P, tmp_path):
    for bad in (dict(nifs=5), dict(nbits=12), dict(nfpc=0), dict(nchans=0)):
        h = _hdr(**bad)
        assert not P.fbh5w_open(C.byref(h), 1, str(tmp_path / "b.h5").encode(), b"", b"", b"")
