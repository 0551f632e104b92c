"""Host <-> device copies (tl_upload / tl_download, tl_snapshot_upload /
tl_snapshot_download): the cells inside both halos move, nothing else is
written, for host halos below, equal to and above the chunk's, on a ragged
mesh.  The copies go through a device staging buffer where they can."""
import ctypes as C

import numpy as np
import pytest

from tealeaf_ref_amd import _lib

pytestmark = pytest.mark.gpu

NX, NY, H = 37, 23, 2


@pytest.fixture(scope="module")
def dev():
    d = _lib.Device(NX, NY, H)
    yield d
    d.close()


def host_array(halo, seed):
    return np.random.default_rng(seed).random((NY + 2 * halo, NX + 2 * halo))


def inner(a, halo, depth):
    """The cells of a host array (halo `halo`) within `depth` of the interior."""
    o = halo - depth
    return a[o:a.shape[0] - o, o:a.shape[1] - o]


@pytest.mark.parametrize("up_halo", [0, 1, 2, 3])
@pytest.mark.parametrize("down_halo", [0, 1, 2, 3])
def test_field_round_trip(dev, up_halo, down_halo):
    print(f"[host copy] field up {up_halo} down {down_halo}", flush=True)
    src = host_array(up_halo, 10 * up_halo + down_halo)
    dev.upload("density", src, halo=up_halo)
    out = np.full((NY + 2 * down_halo, NX + 2 * down_halo), np.nan)
    dev.download("density", halo=down_halo, out=out)
    d = min(up_halo, down_halo, H)
    np.testing.assert_array_equal(inner(out, down_halo, d), inner(src, up_halo, d))
    # a download writes only the cells inside both halos
    k = min(down_halo, H)
    mask = np.ones(out.shape, bool)
    o = down_halo - k
    mask[o:out.shape[0] - o, o:out.shape[1] - o] = False
    assert np.isnan(out[mask]).all()


@pytest.mark.parametrize("halo", [0, 2, 3])
def test_snapshot_round_trip(dev, halo):
    print(f"[host copy] snapshot halo {halo}", flush=True)
    _lib.call("tl_snapshot_reserve", 1)
    src = host_array(halo, 99 + halo)
    ptr = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    _lib.call("tl_snapshot_upload", 0, ptr(src), halo)
    out = np.full(src.shape, np.nan)
    _lib.call("tl_snapshot_download", 0, ptr(out), halo)
    d = min(halo, H)
    np.testing.assert_array_equal(inner(out, halo, d), inner(src, halo, d))
    mask = np.ones(out.shape, bool)
    o = halo - d
    mask[o:out.shape[0] - o, o:out.shape[1] - o] = False
    assert np.isnan(out[mask]).all()
