"""Cross-section handling of the host front-end: the XSL reader's transport
cross section and chi normalization (src/core/material.cpp:57-61,
material_lib.cpp), and replacing a deck's library by a branch case without
re-tracing (mocc_host_reload_xs). CPU only."""
import os

import numpy as np

from mocc_amd import lattice_gen as lg
from mocc_amd.native import HostProblem


def _expected(mats):
    G = len(mats[0].absorption)
    tr, nf, f, ch, sc = [], [], [], [], []
    for m in mats:
        tr.append(m.absorption + m.scatter.sum(axis=0))  # absorption + out-scatter
        nf.append(m.nu_fission)
        f.append(m.fission)
        c = m.chi / m.chi.sum() if m.chi.sum() > 0 else m.chi
        ch.append(c)
        sc.append(m.scatter)
    return (np.concatenate(tr), np.concatenate(nf), np.concatenate(f), np.concatenate(ch),
            np.concatenate([s.ravel() for s in sc]), G)


def _view_xs(hp):
    p = hp.p
    G, M = p.n_group, p.n_mat
    return (hp.array("xs_tr", M * G), hp.array("xs_nf", M * G), hp.array("xs_f", M * G),
            hp.array("xs_ch", M * G), hp.array("xs_sc", M * G * G))


def test_deck_xs_match_materials(deck_dir):
    path = lg.write_config("pins3x3", deck_dir)
    hp = HostProblem.load(path)
    # materials of the deck, in material-id order (FUEL0 id 1, MOD0 id 2)
    tr, nf, f, ch, sc, G = _expected(lg.deck_materials(lg.pins3x3_deck(), 1))
    got = _view_xs(hp)
    for a, b in zip(got, (tr, nf, f, ch, sc)):
        np.testing.assert_allclose(a, b, rtol=1e-15, atol=0)


def test_reload_branch_case(deck_dir):
    path = lg.write_config("pins3x3", deck_dir)
    hp = HostProblem.load(path)
    n_seg, fm = hp.p.n_seg, hp.array("fsr_mat", hp.n_fsr, np.int64)
    xsl = os.path.join(deck_dir, "c1_branch7.xsl")
    mats = lg.write_branch_xsl(lg.pins3x3_deck(), 1, 7, 0.1, xsl)
    base = _view_xs(hp)
    hp.reload_xs(xsl)
    tr, nf, f, ch, sc, G = _expected(mats)
    got = _view_xs(hp)
    for a, b in zip(got, (tr, nf, f, ch, sc)):
        np.testing.assert_allclose(a, b, rtol=1e-15, atol=0)
    assert not np.allclose(got[0], base[0])
    # geometry untouched
    assert hp.p.n_seg == n_seg
    assert np.array_equal(hp.array("fsr_mat", hp.n_fsr, np.int64), fm)


def test_reload_rejects_missing_material(deck_dir, tmp_path):
    import pytest
    path = lg.write_config("pins3x3", deck_dir)
    hp = HostProblem.load(path)
    bad = tmp_path / "bad.xsl"
    mats = lg.deck_materials(lg.pins3x3_deck(), 1)
    mats[0].name = "OTHER"
    lg.write_xsl(str(bad), mats)
    with pytest.raises(RuntimeError, match="not found"):
        hp.reload_xs(str(bad))
