"""Pinning of the oracle's energy functions (ref_field_energy,
ref_uniaxial_energy, ref_zeeman_energy, ref_demag_energy) by closed forms:

  * exchange of a spin wave m = (0, cos kx, sin kx): density A k^2, with the
    P1 interpolation error ~ (kh)^2 / 12, shrinking 4x when h halves;
  * bulk DMI of the same helix: density -D k (surface term m x n . m = 0);
  * uniaxial: 0 for m || k, Ku for m perpendicular to k;
  * Zeeman: -Js H . m for uniform m and H;
  * demag: covered with Aharoni's factors in test_oracle_physics.py and here
    by E = -1/2 sum Js V H.m with H = -Js/(3 mu0) m inside a cube (N = 1/3).
"""
import numpy as np
import pytest

MU0 = 1.256637061435917295376e-6


def _helix(oracle_mod, nx, h):
    from tetmag_amd import mesh as M
    mesh = M.kuhn_cuboid(nx, 4, 4, h, 0.0, 1)
    o = oracle_mod.Oracle(mesh.xyz, mesh.fel)
    n = o.n
    k = 2 * np.pi / (40.0)
    x = mesh.xyz[:, 0]
    m = np.concatenate([np.zeros(n), np.cos(k * x), np.sin(k * x)])
    return o, m, k


def test_exchange_and_dmi_energy_of_a_helix(oracle_mod):
    A_si, D_si = 1.3e-11, 3e-3
    errs_x, errs_d = [], []
    for nx, h in ((40, 1.0), (80, 0.5)):
        o, m, k = _helix(oracle_mod, nx, h)
        n = o.n
        V = o.nodevol.sum()
        f = {"Hexc": o.exchange(m, np.full(n, A_si / 1e-18)), "Hdmi": o.dmi(m, np.full(n, D_si / 1e-9))}
        e = o.energies(m, f, Js=np.ones(n))
        ex = A_si * (k * 1e9) ** 2
        ed = -D_si * k * 1e9
        errs_x.append(abs(e["exchange"] / V - ex) / ex)
        errs_d.append(abs(e["dmi"] / V - ed) / abs(ed))
    assert errs_x[0] < 5e-3 and errs_d[0] < 1e-2
    assert errs_x[1] < errs_x[0] / 3.5  # O(h^2)
    assert errs_d[1] < errs_d[0] / 1.8


def test_uniaxial_and_zeeman_energy(oracle_mod):
    from tetmag_amd import mesh as M
    mesh = M.kuhn_cuboid(4, 3, 2, 2.0, 0.1, 3)
    o = oracle_mod.Oracle(mesh.xyz, mesh.fel)
    n = o.n
    V = o.nodevol.sum()
    k = np.array([0.36, 0.48, 0.8])
    axis = np.concatenate([np.full(n, c) for c in k])
    Ku = np.full(n, 5e4)
    perp = np.array([0.8, 0.0, -0.36]) / np.linalg.norm([0.8, 0.0, -0.36])
    mp = np.concatenate([np.full(n, c) for c in perp])
    f = {"Hexc": np.zeros(3 * n)}
    assert abs(o.energies(axis, f, Js=np.ones(n), Ku=Ku, axis=axis)["uniaxial"]) < 1e-9 * 5e4 * V
    assert o.energies(mp, f, Js=np.ones(n), Ku=Ku, axis=axis)["uniaxial"] / V == pytest.approx(5e4, rel=1e-12)
    Js = np.full(n, 1.2)
    H = np.array([1e4, -2e4, 3e4])
    f["Hext"] = np.concatenate([np.full(n, c) for c in H])
    e = o.energies(mp, f, Js=Js)["zeeman"] / V
    assert e == pytest.approx(-1.2 * H @ perp, rel=1e-12)


def test_demag_energy_formula(oracle_mod):
    from tetmag_amd import mesh as M
    mesh = M.kuhn_cuboid(3, 3, 3, 1.0, 0.0, 1)
    o = oracle_mod.Oracle(mesh.xyz, mesh.fel)
    n = o.n
    Js = np.full(n, 1.0)
    m = np.concatenate([np.zeros(n), np.zeros(n), np.ones(n)])
    H = -m / (3 * MU0)  # the ideal field inside a uniformly magnetised sphere/cube average
    e = o.energies(m, {"Hexc": np.zeros(3 * n), "Hdem": H}, Js=Js)["demag"] / o.nodevol.sum()
    assert e == pytest.approx(1.0 / (6 * MU0), rel=1e-12)  # Js^2 / (6 mu0) = 1/2 mu0 Ms^2 / 3
