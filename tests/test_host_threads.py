"""The host tracer and the array assembly run over the azimuths on several
threads (MOCC_HOST_THREADS). Each azimuth is traced and copied on its own, so
every traced array must be the same, bit for bit, at any thread count. CPU
only."""
import numpy as np
import pytest

from mocc_amd import lattice_gen as lg
from mocc_amd.native import HostProblem


def _arrays(hp):
    p = hp.p
    n_ang = p.n_polar * p.n_azi
    sizes = dict(ang_weight=n_ang, ang_rsin_theta=n_ang, ang_alpha=n_ang,
                 azi_ray_off=p.n_azi + 1, azi_nx=p.n_azi, azi_ny=p.n_azi, azi_spacing=p.n_azi,
                 ray_seg_off=p.n_ray + 1, ray_bc=2 * p.n_ray, seg_fsr=p.n_seg, seg_len=p.n_seg,
                 ray_cm_cell=2 * p.n_ray, ray_cm_surf=2 * p.n_ray, ray_cm_off=p.n_ray + 1,
                 fsr_vol=p.n_fsr, fsr_area=p.n_fsr, fsr_mat=p.n_fsr)
    out = {k: np.ctypeslib.as_array(getattr(p, k), shape=(n,)).copy() for k, n in sizes.items()}
    out["cm_data"] = np.ctypeslib.as_array(p.cm_data, shape=(4 * int(out["ray_cm_off"][-1]),)).copy()
    return out


@pytest.mark.parametrize("deck", ["lat12_vacuum", "assembly17_coarse"])
def test_trace_independent_of_thread_count(deck_dir, monkeypatch, deck):
    if deck == "lat12_vacuum":
        path = lg.write_deck(lg.pin_lattice(12, 2, 8, 0.05, "thr12", cmfd=True,
                                            bc={"east": "vacuum", "south": "vacuum"}), deck_dir, 12)
    else:
        path = lg.write_config("assembly17", deck_dir, name="thr17", n_azimuthal=8, spacing=0.05)
    ref = None
    for n in ("1", "3", "8"):
        monkeypatch.setenv("MOCC_HOST_THREADS", n)
        got = _arrays(HostProblem.load(path))
        if ref is None:
            ref = got
            assert got["seg_len"].size > 0
            continue
        for k in ref:
            assert np.array_equal(ref[k], got[k]), f"{k} differs at {n} threads"
