"""Host front-end (quadrature, ray tracer, RayData) against the reference's own
unit-test values. CPU only.

Golden values are the ones asserted by the reference's tests:
  src/core/tests/test_AngularQuadrature.cpp:182-256
  src/sweepers/moc/tests/test_Ray.cpp:43-141, 143-305
  src/sweepers/moc/tests/test_RayData.cpp:31-72
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

from conftest import DECKS
from mocc_amd.native import HostProblem, host

PI = math.pi


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def quad(xml):
    out = np.zeros(8 * 512 * 7)
    n = host().mocc_host_quadrature_xml(xml.encode(), _p(out), 8 * 512)
    assert n > 0, host().mocc_host_last_error()
    return out[: 7 * n].reshape(n, 7)


def mesh(xml):
    h = C.c_void_p()
    rc = host().mocc_host_load_mesh_string(xml.encode(), DECKS.encode(), C.byref(h))
    assert rc == 0, host().mocc_host_last_error()
    return h


def trace(h, x1, y1, x2, y2, cap=4096):
    idx = np.zeros(cap, np.int32)
    ln = np.zeros(cap)
    head = np.zeros(5, np.int32)
    cm = np.zeros(4 * cap, np.int32)
    d = C.c_double
    n = host().mocc_host_trace_ray(h, d(x1), d(y1), d(x2), d(y2), _p(idx), _p(ln), cap, _p(head), _p(cm), cap)
    assert n >= 0, host().mocc_host_last_error()
    return idx[:n], ln[:n], head, cm[: 4 * head[4]].reshape(-1, 4)


# --- angular quadrature ---------------------------------------------------------

def test_cy_16_3():
    q = quad('<ang_quad type="cy" n_azimuthal="16" n_polar="3" />')
    assert q.shape[0] == 8 * 48
    assert abs(q[:, 5].sum() - 8.0) < 1e-14
    gold = [0.049087385212340, 0.167429147795000, 0.166447265186000, 0.008177029791330,
            0.986016452244020, 0.002889562500000, 6.000672075260800]
    np.testing.assert_allclose(q[0], gold, rtol=0, atol=1e-13)


def test_cg_16_3():
    q = quad('<ang_quad type="cg" n_azimuthal="16" n_polar="3" />')
    assert q.shape[0] == 8 * 48
    assert abs(q[:, 5].sum() - 8.0) < 1e-13
    gold = [0.049087385212340, 1.329852612388110, 0.9699434705550130, 0.0476502671628313,
            0.2386191860831970, 0.029244620910793, 1.029746048636890]
    np.testing.assert_allclose(q[0], gold, rtol=0, atol=1e-12)


@pytest.mark.parametrize("order,ndir", [(2, 1), (4, 3), (6, 6), (8, 10), (12, 21), (16, 36)])
def test_level_symmetric(order, ndir):
    q = quad(f'<ang_quad type="ls" order="{order}" />')
    assert q.shape[0] == 8 * ndir
    assert abs(q[:, 5].sum() - 8.0) < 1e-14
    # every direction is a unit vector
    np.testing.assert_allclose(q[:, 2] ** 2 + q[:, 3] ** 2 + q[:, 4] ** 2, 1.0, atol=1e-14)


def test_octant_expansion_and_reverse():
    # octant order azimuth-major, polar-minor; reverse(i) = (i + 2 n_oct) mod 4 n_oct
    # (angular_quadrature.hpp:158-167) flips the in-plane direction
    q = quad('<ang_quad type="cy" n_azimuthal="4" n_polar="3" />')
    n = 12
    for i in range(4 * n):
        r = (i + 2 * n) % (4 * n)
        np.testing.assert_allclose(q[r, 2:4], -q[i, 2:4], atol=1e-14)
        np.testing.assert_allclose(q[r, 4], q[i, 4], atol=1e-14)
    # octants 1-4 have +z, 5-8 the mirrored -z
    np.testing.assert_allclose(q[4 * n:, 4], -q[: 4 * n, 4], atol=1e-14)


def test_cg_3_1_update_weights():
    out = np.zeros(8 * 3 * 7)
    idx = np.array([0, 2], np.int32)
    alpha = np.array([PI / 16.0, 7.0 * PI / 16.0])
    n = host().mocc_host_quadrature_modify(b'<ang_quad type="cg" n_azimuthal="3" n_polar="1" />', 2,
                                            _p(idx), _p(alpha), _p(out), 8 * 3)
    assert n == 24, host().mocc_host_last_error()
    q = out.reshape(n, 7)
    assert abs(q[1, 0] - PI / 4.0) < 1e-12
    np.testing.assert_allclose(q[:3, 5], [0.3125, 0.375, 0.3125], rtol=0, atol=1e-14)


# --- rays -----------------------------------------------------------------------

WEIRD_IDX = [29, 28, 27, 26, 25, 105, 104, 103, 102, 101, 100, 180, 179, 178, 177, 176, 175, 255,
             254, 253, 252, 251, 250, 330, 329, 328, 327, 326, 325, 405, 404, 403, 402, 401, 400, 480]
WEIRD_LEN = [
    0.12752525252525659, 0.12752525252525249, 0.12752525252525249, 0.12752525252525268,
    0.12114898989898612, 0.0063762626262663788, 0.12752525252525249, 0.12752525252525249,
    0.12752525252525249, 0.12752525252525249, 0.11477272727272445, 0.012752525252528228,
    0.12752525252525249, 0.12752525252525249, 0.12752525252525249, 0.12752525252525249,
    0.10839646464646216, 0.01912878787879034, 0.12752525252525268, 0.12752525252525249,
    0.12752525252525249, 0.12752525252525249, 0.10202020202019987, 0.025505050505052623,
    0.12752525252525268, 0.12752525252525249, 0.12752525252525249, 0.12752525252525249,
    0.095643939393937588, 0.031881313131314912, 0.12752525252525249, 0.12752525252525249,
    0.12752525252525268, 0.12752525252525249, 0.089267676767675302, 0.038257575757577197]


def test_weird_ray():
    ones = " ".join(["1"] * 6400)
    xml = f"""<mesh id="1" type="rect" pitch="10"><sub_x>80</sub_x><sub_y>80</sub_y></mesh>
<pin id="1" mesh="1">{ones}</pin>
<lattice id="1" nx="1" ny="1">1</lattice>
<assembly id="1" np="1" hz="1.0"><lattices>1</lattices></assembly>
<core nx="1" ny="1" north="reflect" south="reflect" top="reflect" bottom="reflect" west="prescribed" east="prescribed">1</core>
<material_lib path="c5g7.xsl"><material id="1" name="UO2-3.3" /></material_lib>"""
    h = mesh(xml)
    try:
        idx, ln, _, _ = trace(h, -1.249999999999996 + 5.0, -5.0 + 5.0, -5 + 5.0, -4.2424242424242422 + 5.0)
        assert idx.tolist() == WEIRD_IDX
        np.testing.assert_allclose(ln, WEIRD_LEN, rtol=0, atol=1e-15)
    finally:
        host().mocc_host_free(h)


EAST, NORTH, WEST, SOUTH = 0, 1, 2, 3


def test_simple_ray_coarse_data():
    h = mesh(open(os.path.join(DECKS, "6x5.xml")).read())
    try:
        idx, ln, head, cm = trace(h, 0.0, 1.0, 4.0, 5.0)
        cell_fw, surf_fw, cell_bw, surf_bw, ncseg = head
        assert (surf_fw, cell_fw, surf_bw, cell_bw) == (37, 6, 88, 27)
        assert len(idx) == 12 and ncseg == 8
        np.testing.assert_allclose(ln, math.sqrt(2.0) / 3.0, atol=1e-5)
        assert cm[:, 0].tolist() == [EAST, NORTH] * 4
        assert cm[:, 1].tolist() == [WEST, SOUTH, WEST, SOUTH, WEST, SOUTH, SOUTH, WEST]
        assert cm[:, 2].tolist() == [3, 0, 3, 0, 3, 0, 3, 0]
        assert cm[:, 3].tolist() == [3, 0, 3, 0, 3, 0, 3, 0]

        idx, ln, head, cm = trace(h, 4.0, 0.0, 6.0, 2.0)
        cell_fw, surf_fw, cell_bw, surf_bw, ncseg = head
        assert (surf_fw, cell_fw, surf_bw, cell_bw) == (89, 4, 43, 11)
        assert len(idx) == 6 and ncseg == 4
        # corner-crossing rays trace without error
        for r in [(2.0, 0.0, 0.0, 2.0), (6.0, 3.0, 4.0, 5.0), (0.0, 0.5, 6.0, 3.25)]:
            trace(h, *r)
    finally:
        host().mocc_host_free(h)


def test_nasty_rays():
    h = mesh(open(os.path.join(DECKS, "square.xml")).read())
    try:
        for r in [(1.26, 0.0, 3.78, 2.52), (1.26, 0.0, 0.0, 1.26), (0.0, 1.26, 2.52, 3.78),
                  (3.78, 2.52, 2.52, 3.78)]:
            idx, ln, _, _ = trace(h, *r)
            assert len(idx) > 0 and (ln > 0).all()
    finally:
        host().mocc_host_free(h)


# --- RayData --------------------------------------------------------------------

def _deck(geom_file, sweeper):
    return open(os.path.join(DECKS, geom_file)).read() + sweeper


def _ray_volumes(hp):
    """sum over angles and rays of len * spacing * w * 2 pi per FSR, from the
    flattened problem (polar angles share an azimuth's rays)."""
    p = hp.p
    A, Pn, F = p.n_azi, p.n_polar, p.n_fsr
    w = hp.array("ang_weight", A * Pn).reshape(A, Pn)
    sp = hp.array("azi_spacing", A)
    aro = hp.array("azi_ray_off", A + 1, np.int64)
    rso = hp.array("ray_seg_off", p.n_ray + 1, np.int64)
    fsr = hp.array("seg_fsr", p.n_seg, np.int64)
    ln = hp.array("seg_len", p.n_seg)
    vol = np.zeros(F)
    for a in range(A):
        s0, s1 = rso[aro[a]], rso[aro[a + 1]]
        wt = sp[a] * w[a].sum() * 2.0 * PI
        np.add.at(vol, fsr[s0:s1], ln[s0:s1] * wt)
    return vol


def test_raydata_volume_identity():
    hp = HostProblem.from_string(
        _deck("square.xml", '<sweeper type="moc" n_inner="1"><ang_quad type="ls" order="4" /><rays spacing="0.01" /></sweeper>'),
        DECKS)
    try:
        vol = _ray_volumes(hp)
        np.testing.assert_allclose(vol, 0.1764 * 4.0 * PI, rtol=0, atol=1e-14)
    finally:
        hp.close()


def test_raydata_volume_identity_cyl(pins3x3):
    # FLAT correction (ray_data.cpp:350-416): every FSR's ray-traced volume
    # equals its true area, per angle, for the cylindrical pins of a lattice
    areas = np.zeros(pins3x3.n_fsr)
    n = host().mocc_host_fsr_areas(pins3x3.h, _p(areas), pins3x3.n_fsr)
    assert n == pins3x3.n_fsr
    vol = _ray_volumes(pins3x3)
    np.testing.assert_allclose(vol, areas * 4.0 * PI, rtol=1e-12)
