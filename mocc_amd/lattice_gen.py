"""Seeded synthetic mocc input decks (XML + XSL) for the benchmark lattices.

Every deck is a plain mocc input: the same files feed the device path and the
FP64 oracle, so both read bit-identical geometry and cross sections (XSL
values are written with %.17g).

Cross sections per material (draws in material-major, group-minor order):
  * Sigma_t ~ U[0.2, 1.5] /cm, scattering ratio c ~ U[0.5, 0.95];
  * within-group share of the scattering ~ U[0.6, 0.9], the rest down-scatter
    spread over the lower-energy groups with halving weights;
  * 7 groups: groups 5-7 also up-scatter at most 5 % of their scattering into
    the next higher group;
  * absorption = (1 - c) Sigma_t, so that mocc's Sigma_t = absorption +
    out-scatter (src/core/material.cpp:57-61) reproduces the draw;
  * fuel only: nu Sigma_f / Sigma_t ~ U[0.02, 0.1], Sigma_f = nu Sigma_f / 2.43,
    chi = [1, 0] for 2 groups and the C5G7 UO2 spectrum for 7 groups.

Pins are 1.26 cm pitch cylinders of radius 0.54 cm: 3 equal-area fuel rings and
the moderator, each cut into 8 sectors (32 FSRs per pin).
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

C5G7_UO2_CHI = [5.8791e-01, 4.1176e-01, 3.3906e-04, 1.1761e-07, 0.0, 0.0, 0.0]

# 17x17 guide-tube positions (24 guide tubes + the instrument tube), the layout
# of mocc's geometry input documentation (doc/geometry_input.dox:85-103).
GT17 = [(2, 5), (2, 8), (2, 11), (3, 3), (3, 13), (5, 2), (5, 5), (5, 8), (5, 11), (5, 14),
        (8, 2), (8, 5), (8, 8), (8, 11), (8, 14), (11, 2), (11, 5), (11, 8), (11, 11), (11, 14),
        (13, 3), (13, 13), (14, 5), (14, 8), (14, 11)]


@dataclass
class Material:
    name: str
    absorption: np.ndarray
    nu_fission: np.ndarray
    fission: np.ndarray
    chi: np.ndarray
    scatter: np.ndarray  # [to][from]


def make_material(rng: np.random.Generator, name: str, G: int, fissile: bool) -> Material:
    st = rng.uniform(0.2, 1.5, G)
    c = rng.uniform(0.5, 0.95, G)
    within = rng.uniform(0.6, 0.9, G)
    nf_ratio = rng.uniform(0.02, 0.1, G)
    up_frac = rng.uniform(0.0, 0.05, G)
    scat = np.zeros((G, G))
    for g in range(G):
        s_out = c[g] * st[g]
        up = s_out * up_frac[g] if (G == 7 and g >= 4) else 0.0
        if g == G - 1:
            self_s = s_out - up
            down = 0.0
        else:
            self_s = within[g] * s_out
            down = s_out - self_s - up
        scat[g, g] = self_s
        if down > 0.0:
            w = 0.5 ** np.arange(1, G - g)
            w = w / w.sum()
            for k, to in enumerate(range(g + 1, G)):
                scat[to, g] += down * w[k]
        if up > 0.0:
            scat[g - 1, g] += up
    absorption = (1.0 - c) * st
    if fissile:
        nuf = nf_ratio * st
        fis = nuf / 2.43
        chi = np.array([1.0, 0.0]) if G == 2 else np.array(C5G7_UO2_CHI[:G])
    else:
        nuf = np.zeros(G)
        fis = np.zeros(G)
        chi = np.zeros(G)
    return Material(name, absorption, nuf, fis, chi, scat)


def perturb(m: Material, rng: np.random.Generator, rel: float, name: str) -> Material:
    """Branch case: every cross section of the material times U[1-rel, 1+rel]."""
    G = len(m.absorption)
    f = rng.uniform(1 - rel, 1 + rel, 4 + G * G)
    return Material(name, m.absorption * f[0], m.nu_fission * f[1], m.fission * f[1],
                    m.chi.copy(), m.scatter * f[4:].reshape(G, G))


def write_xsl(path: str, mats: list[Material]) -> None:
    G = len(mats[0].absorption)
    bounds = np.logspace(7.3, -1.2, G)
    lines = ["Synthetic seeded macroscopic cross sections", f" {G} {len(mats)}",
             " " + " ".join("%.6E" % b for b in bounds)]
    for m in mats:
        lines.append(f"XSMACRO {m.name} 0")
        for g in range(G):
            lines.append("  " + " ".join("%.17g" % v for v in
                                         (m.absorption[g], m.nu_fission[g], m.fission[g], m.chi[g])))
        for to in range(G):
            lines.append("  " + " ".join("%.17g" % v for v in m.scatter[to]))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


@dataclass
class Deck:
    """Geometry description: a core of lattices of pins (pin ids 1 = fuel of
    material set s -> pin id 10*s+1, guide tube -> 10*s+2)."""
    name: str
    groups: int
    n_azimuthal: int
    spacing: float
    lattices: dict = field(default_factory=dict)  # id -> (nx, ny, rows of pin ids, top row first)
    core: list = field(default_factory=list)      # rows of lattice ids, top row first
    bc: dict = field(default_factory=dict)
    n_sets: int = 1
    cmfd: bool = False
    n_inner: int = 1
    k_tol: float = 1e-6
    psi_tol: float = 1e-5
    max_iter: int = 500


def lattice_rows(n: int, set_id: int, guide_tubes: bool) -> list[list[int]]:
    fuel, gt = 10 * set_id + 1, 10 * set_id + 2
    rows = [[fuel] * n for _ in range(n)]
    if guide_tubes:
        for (i, j) in GT17:
            rows[i][j] = gt
    return rows


def deck_materials(d: Deck, seed: int) -> list[Material]:
    """The deck's base materials (FUEL<s>, MOD<s> per material set)."""
    rng = np.random.default_rng(seed)
    mats = []
    for s in range(d.n_sets):
        mats.append(make_material(rng, f"FUEL{s}", d.groups, True))
        mats.append(make_material(rng, f"MOD{s}", d.groups, False))
    return mats


def write_branch_xsl(d: Deck, seed: int, case_seed: int, rel: float, path: str) -> list[Material]:
    """A cross-section branch case of the deck: every base cross section times
    U[1-rel, 1+rel] drawn from case_seed; same material names, so it can
    replace the deck's library (mocc_host_reload_xs)."""
    rng = np.random.default_rng(case_seed)
    mats = [perturb(m, rng, rel, m.name) for m in deck_materials(d, seed)]
    write_xsl(path, mats)
    return mats


def write_deck(d: Deck, out_dir: str, seed: int, rel_perturb: float = 0.0) -> str:
    """Write <out_dir>/<name>.xml and .xsl; return the XML path."""
    os.makedirs(out_dir, exist_ok=True)
    rng = np.random.default_rng(seed)
    mats = []
    for s in range(d.n_sets):
        mats.append(make_material(rng, f"FUEL{s}", d.groups, True))
        mats.append(make_material(rng, f"MOD{s}", d.groups, False))
    if rel_perturb > 0.0:
        mats = [perturb(m, rng, rel_perturb, m.name) for m in mats]
    xsl = f"{d.name}.xsl"
    write_xsl(os.path.join(out_dir, xsl), mats)
    bc = {"north": "reflect", "south": "reflect", "east": "reflect", "west": "reflect"}
    bc.update(d.bc)
    x = []
    x.append(f'<solver type="eigenvalue" k_tol="{d.k_tol:g}" psi_tol="{d.psi_tol:g}" '
             f'max_iter="{d.max_iter}" cmfd="{"t" if d.cmfd else "f"}">')
    x.append('  <source scattering="P0" />')
    x.append(f'  <sweeper type="moc" n_inner="{d.n_inner}">')
    x.append(f'    <ang_quad type="cy" n_azimuthal="{d.n_azimuthal}" n_polar="3" />')
    x.append(f'    <rays spacing="{d.spacing:.17g}" />')
    x.append('  </sweeper>')
    if d.cmfd:
        x.append('  <cmfd enabled="t" />')
    x.append('</solver>')
    x.append(f'<material_lib path="{xsl}">')
    for s in range(d.n_sets):
        x.append(f'  <material id="{10 * s + 1}" name="FUEL{s}" />')
        x.append(f'  <material id="{10 * s + 2}" name="MOD{s}" />')
    x.append('</material_lib>')
    x.append('<mesh id="1" type="cyl" pitch="1.26">')
    x.append('  <radii>0.54</radii><sub_radii>3</sub_radii><sub_azi>8</sub_azi>')
    x.append('</mesh>')
    for s in range(d.n_sets):
        x.append(f'<pin id="{10 * s + 1}" mesh="1">{10 * s + 1} {10 * s + 2}</pin>')
        x.append(f'<pin id="{10 * s + 2}" mesh="1">{10 * s + 2} {10 * s + 2}</pin>')
    for lid, (nx, ny, rows) in d.lattices.items():
        x.append(f'<lattice id="{lid}" nx="{nx}" ny="{ny}">')
        for r in rows:
            x.append("  " + " ".join(str(v) for v in r))
        x.append('</lattice>')
        x.append(f'<assembly id="{lid}" np="1" hz="1.0"><lattices>{lid}</lattices></assembly>')
    ny_c, nx_c = len(d.core), len(d.core[0])
    x.append(f'<core nx="{nx_c}" ny="{ny_c}" north="{bc["north"]}" south="{bc["south"]}" '
             f'east="{bc["east"]}" west="{bc["west"]}" top="reflect" bottom="reflect">')
    for r in d.core:
        x.append("  " + " ".join(str(v) for v in r))
    x.append('</core>')
    path = os.path.join(out_dir, f"{d.name}.xml")
    with open(path, "w") as f:
        f.write("\n".join(x) + "\n")
    return path


# --- the benchmark configurations ---------------------------------------------

def pin_lattice(n: int, groups: int, n_azimuthal: int, spacing: float, name: str,
                guide_tubes: bool = False, **kw) -> Deck:
    d = Deck(name, groups, n_azimuthal, spacing, **kw)
    d.lattices = {1: (n, n, lattice_rows(n, 0, guide_tubes))}
    d.core = [[1]]
    return d


def pins3x3_deck() -> Deck:
    """3x3 pins, 2 groups, CY 16x3, spacing 0.05 cm, all reflective."""
    return pin_lattice(3, 2, 16, 0.05, "pin3x3_2g")


def assembly17_deck() -> Deck:
    """17x17 assembly with 25 moderator-filled guide tubes, 7 groups, CY 32x3, 0.03 cm."""
    return pin_lattice(17, 7, 32, 0.03, "asm17_7g", guide_tubes=True)


def core51_deck() -> Deck:
    """3x3 assemblies of 17x17 (3 material sets), 7 groups, CY 64x3, 0.02 cm,
    vacuum east/south, reflective west/north, CMFD on, 5 inner iterations
    (as in mocc's example inputs: with a single inner iteration the D-hat of
    the first, unconverged sweep makes the coarse matrix lose its positive
    diagonal on this deck and mocc's CMFD iteration diverges)."""
    d = Deck("core3x3_7g", 7, 64, 0.02, n_sets=3, cmfd=True, n_inner=5,
             bc={"east": "vacuum", "south": "vacuum"})
    d.lattices = {s + 1: (17, 17, lattice_rows(17, s, True)) for s in range(3)}
    d.core = [[1, 2, 3], [2, 3, 1], [3, 1, 2]]
    return d


CONFIGS = {"pins3x3": (pins3x3_deck, 1), "assembly17": (assembly17_deck, 2), "core51": (core51_deck, 3)}


def write_config(key: str, out_dir: str, **overrides) -> str:
    fn, seed = CONFIGS[key]
    d = fn()
    for k, v in overrides.items():
        setattr(d, k, v)
    return write_deck(d, out_dir, seed)
