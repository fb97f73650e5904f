"""Pin the CPU oracle (oracle/fzcpu.c) to the reference's own outputs.

The golden fixtures were written by the reference's compiled Fortran
(tests/golden/make_golden.py); where the reference build (oracle/_ref) is
present the oracle is also compared with it directly on fresh inputs.
"""
import numpy as np
import pytest

from firtez_dz_amd import lines

CART = (1, 1, 1, 1, 1, 1, 1, 0)


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


def test_partition_functions_bitexact(cpu, golden):
    g = golden["units"]
    for i in range(92):
        for j, t in enumerate(g["pf_T"]):
            u, du = cpu.partition(i + 1, float(t))
            assert np.array_equal(u, g["pf_U"][i, j]), (i + 1, t)
            assert np.array_equal(du, g["pf_DU"][i, j]), (i + 1, t)


def test_eos_and_opacities(cpu, golden):
    g = golden["units"]
    cpu.setup(lines.fe6301_6302(), resp=True, atmpar=CART)
    for i in range(len(g["eos_T"])):
        T, P = float(g["eos_T"][i]), float(g["eos_P"][i])
        e, out, dv = cpu.eos(T, P, 0.0)
        assert e == 0
        assert rel(out, g["eos_out"][i]) < 1e-13
        # DERIVVAR written by the EOS: d n_H and d n_e derivatives
        for k in (0, 1, 2, 3, 8, 9, 10, 12):
            assert abs(dv[k] - g["eos_dv"][i, k]) <= 1e-12 * abs(g["eos_dv"][i, k]) + 1e-300, k
        nh, ne = out[0], out[1]
        for j, lam in enumerate((5000.0, lines.fe6301_6302().lines[0].lambda0)):
            _, kc, d = cpu.opac_cont(T, lam, ne, nh, dv)
            assert rel(kc, g["kc"][i, j]) < 1e-13
            assert rel(d[17:21], g["dkc"][i, j]) < 1e-11
        for j in range(2):
            _, kl, d = cpu.opac_line(j, T, ne, nh, dv)
            assert rel(kl, g["kl"][i, j]) < 1e-13
            assert rel(d[[13, 14, 15, 16, 21, 22]], g["dkl"][i, j]) < 1e-11
            dm, d2 = cpu.damping(j, T, nh, d)
            assert rel(dm, g["damp"][i, j]) < 1e-13
            assert rel(d2[4:8], g["ddamp"][i, j]) < 1e-11


def test_voigt_and_zeeman(cpu, golden):
    g = golden["units"]
    for i, a in enumerate(g["voigt_a"]):
        h, f = cpu.voigt(float(a), g["voigt_v"])
        assert np.allclose(h, g["voigt_H"][i], rtol=1e-13, atol=1e-15)
        assert np.allclose(f, g["voigt_F"][i], rtol=1e-13, atol=1e-15)
    cpu.setup(lines.fe6301_6302(), resp=False)
    for j in range(2):
        s, st, ty = cpu.zeeman(j)
        assert np.array_equal(s, g["zeeman%d_shift" % j])
        assert np.allclose(st, g["zeeman%d_strength" % j], rtol=1e-15)
        assert np.array_equal(ty, g["zeeman%d_type" % j])


def test_syn1_stokes_and_gain(cpu, golden):
    g = golden["syn1"]
    ls = lines.setup([((4, 5), 500, -700.0, 5.0)])
    cpu.setup(ls, resp=False)
    gain = cpu.gain()
    # GAIN1D agrees to one REAL(4) ulp (a rounding boundary on a few pixels)
    assert np.max(np.abs(gain - g["gain"]) / np.abs(g["gain"])) < 2.5e-7
    cpu.set_gain(g["gain"])
    mo, syn, _, kc5 = cpu.forward1d(g["model_in"])
    assert np.max(np.abs(syn - g["stokes"])) < 1e-12
    assert rel(kc5, g["kc5"]) < 1e-12
    assert np.allclose(mo, g["model_out"], rtol=2e-7, atol=0)


def test_rf_test_atmosphere(cpu, golden):
    g = golden["rf"]
    ls = lines.setup([((4, 5), 500, -700.0, 5.0)])
    cpu.setup(ls, resp=True, atmpar=CART)
    cpu.set_gain(g["gain"])
    mo, syn, dsyn, _ = cpu.forward1d(g["model_in"])
    assert np.max(np.abs(syn - g["stokes"])) < 1e-12
    ours = dsyn[g["rf_heights"]][:, :, :7, :]
    for p in range(7):
        ref = g["rf"][:, :, p, :]
        scale = np.abs(ref).max()
        assert np.max(np.abs(ours[:, :, p, :] - ref)) <= 1e-6 * scale + 1e-300, p


def test_seeded_columns(cpu, golden):
    g = golden["seeded_c1"]
    cpu.setup(lines.fe6301_6302(), resp=True, atmpar=CART)
    cpu.set_gain(g["gain"])
    for i in range(len(g["cols"])):
        m = g["models"][i]
        mo, syn, dsyn, _ = cpu.forward1d(m)
        assert np.max(np.abs(syn - g["stokes"][i])) < 1e-12
        if i < len(g["rf"]):
            for p in range(7):
                ref = g["rf"][i][:, :, p, :]
                assert np.max(np.abs(dsyn[:, :, p, :] - ref)) <= 1e-6 * np.abs(ref).max() + 1e-300


def test_hydrostatic_matches_fixture(cpu, golden):
    g = golden["seeded_c1"]
    for i in range(3):
        m = g["models"][i]
        pg, rho, pel, mw, _ = cpu.he(m[0], float(m[1, -1]), g["z"])
        assert np.array_equal(pg, m[1])
        assert np.array_equal(rho, m[2])


@pytest.mark.skipif("not __import__('oracle').ref_available()")
def test_oracle_vs_reference_build(cpu):
    """Fresh random inputs through both the restatement and the reference build."""
    import oracle
    r = oracle.Ref()
    rng = np.random.default_rng(11)
    ls = lines.fe6301_6302(nwave=64, start=-500.0, step=25.0)
    cpu.setup(ls, resp=True, atmpar=CART)
    r.setup(ls, resp=True, atmpar=CART)
    g = r.gain()
    cpu.set_gain(g)
    r.set_gain(g)
    nz = 48
    z = np.arange(nz) * 20.0
    for _ in range(3):
        T = (4200 + 4500 * np.exp(-z / 180.0) + rng.normal(0, 150, nz)).astype(np.float32)
        pg, rho, pel, mw, _ = r.he(T, 30.0, z)
        m = np.zeros((13, nz), np.float32)
        m[0], m[1], m[2], m[3], m[4] = T, pg, rho, pel, mw
        m[5:8] = rng.uniform(-1500, 1500, 3)[:, None]
        m[10] = rng.normal(0, 2e5)
        m[12] = z
        a = cpu.forward1d(m)
        b = r.forward1d(m)
        assert np.max(np.abs(a[1] - b[1])) < 1e-12
        for p in range(7):
            assert np.max(np.abs(a[2][:, :, p] - b[2][:, :, p])) <= 1e-9 * np.abs(b[2][:, :, p]).max()


def test_hydrostatic_mode_vs_fixture(cpu, golden):
    """HYDROSTATIC: yes: the RK4 re-integration inside the forward, the
    hydrostatic term of dI/dT, P0, and the reference's use of the outgoing
    Stokes vector for the rho, B and v responses in that mode."""
    g = golden["he"]
    ls = lines.setup([((4, 5), int(g["nwave"]), float(g["wave_start"]), float(g["wave_step"]))])
    cpu.setup(ls, resp=True, atmpar=(1, 1, 1, 1, 1, 1, 1, 1))
    cpu.set_gain(g["gain"])
    cpu.set_hydrostatic(True)
    try:
        mo, syn, dsyn, _ = cpu.forward1d(g["model_in"])
    finally:
        cpu.set_hydrostatic(False)
    for row in (1, 2, 3, 4, 11):
        assert np.array_equal(mo[row], g["model_out"][row]), row
    assert np.max(np.abs(syn - g["stokes"])) < 1e-12
    ours = dsyn[g["rf_heights"]]
    for p in range(8):
        sc = np.abs(g["rf"][:, :, p, :]).max()
        assert sc > 0, p
        assert np.max(np.abs(ours[:, :, p, :] - g["rf"][:, :, p, :])) <= 1e-6 * sc, p
