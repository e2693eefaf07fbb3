"""Host setup: Settings.ini -> problem, pinned against reference-held data
(TAU_Line.txt, RC_Line.txt, the shipped Settings.ini files) and closed forms."""

from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np
import pytest

import fdtd2d_amd as F

from helpers import CJ, EPS0, MU0

GOLD = Path(__file__).resolve().parent / "golden"
INIS = ["settings_full.ini", "settings_release.ini", "settings_build.ini"]
C0 = 2.99792458e8
ETA0 = math.sqrt(MU0 / EPS0)


@pytest.fixture(scope="module")
def release():
    return F.Setup.from_ini(str(GOLD / "settings_release.ini"))


def test_dt_matches_tau_line_fixture(release):
    """TAU_Line.txt holds (steps, picoseconds) pairs: their ratio is dt of a
    ds = 0.5 mm, dt_coef = 0.5 run under the 3D Courant rule."""
    rows = json.loads((GOLD / "reference_lines.json").read_text())["tau_line"]
    dt_ps = release.get("dt") * 1e12
    for steps, ps in rows:
        assert ps / steps == pytest.approx(dt_ps, rel=2e-5)
    assert dt_ps == pytest.approx(0.99 / (C0 * math.sqrt(3 / 0.5e-3 ** 2)) * 0.5 * 1e12, rel=1e-12)


@pytest.mark.parametrize("name", INIS)
def test_shipped_settings_build(name):
    s = F.Setup.from_ini(str(GOLD / name))
    assert s.ini_status == 0
    p = s.problem()
    assert p.npml == 11 and p.medium_kind == F.MEDIUM_TABLE
    # Use_Exp_time = 1: 16 ns / dt, truncated
    assert p.nsteps == int(16e-9 / p.dt)
    assert p.nx == int(s.get("Imax")) and p.ny == 3000


def test_array_positions_follow_reference_index_rules(release):
    """calculate2Darrai: qantI = 80/10 + 1 elements from Isource - 40 cells
    with pitch 10; the kernel adds the element index to the column and the
    K offset 1 to the row (F/kernel.cu:641-656, 782-802)."""
    p = release.problem()
    e = np.arange(9)
    assert list(p.src_i) == list(210 + 10 * e + e)
    assert set(p.src_j) == {100 + 1}
    assert list(p.det_i) == list(210 + 10 * e + e)
    assert set(p.det_j) == {200 + 1}
    # component detector: source columns on the detector row
    assert list(p.comp_i) == list(p.src_i) and set(p.comp_j) == {201}


def test_source_profile_and_table(release):
    p = release.problem()
    amp = release.get("amp")
    front, plato = int(release.get("Front")), int(release.get("Plato"))
    fw = front // 2
    tab = p.src_table.astype(np.float64)
    assert np.all(tab[0] == 0) and np.all(tab[-1] == 0)  # n = 0 and n = nMax read 0
    # envelope: Gaussian rise over fw steps, plateau at amp/9, fall
    env = np.max(np.abs(tab[:, 0].reshape(-1)[1: 2 * fw + plato]))
    assert env == pytest.approx(amp / 9, rel=1e-6)
    n = fw // 2
    g = amp * math.exp(-((n - fw) ** 2) / (fw / 2.5) ** 2) / 9
    assert abs(tab[n, 0]) <= g * (1 + 1e-5)
    # carrier: cos(2 * 3.14 * f0 t) - the reference's 3.14 puts the line at
    # FreqA * 3.14 / pi, not FreqA (zero-padded FFT for sub-bin resolution)
    seg = tab[fw: fw + plato, 0]
    nfft = 32 * len(seg)
    spec = np.abs(np.fft.rfft(seg * np.hanning(len(seg)), nfft))
    f = np.fft.rfftfreq(nfft, p.dt)[np.argmax(spec)]
    assert f == pytest.approx(38e9 * 3.14 / math.pi, rel=3e-4)
    # SPAngle = 0: every element carries the same value
    assert np.array_equal(tab[:, 0], tab[:, 5])


def test_source_table_arithmetic_matches_reference_expression(release):
    """buildSourceProfile's double Gaussian front stored as float, then
    Dev_AESASourceCalc / Dev_calculateChirp: float t = n*dt, float phase,
    double 2*3.14 and double cos, float amplitude / element count
    (F/FDTD_cuda.cu:3565-3597, F/kernel.cu:654, 694-698), redone in numpy."""
    p = release.problem()
    f32 = np.float32
    amp, fw = release.get("amp"), int(release.get("Front")) // 2
    dt, f0 = f32(p.dt), f32(release.get("fCoef"))
    ns = np.arange(1, 300)
    env = np.array([f32(amp * math.exp(-1.0 * (n - fw) ** 2 / (fw / 2.5) ** 2)) for n in ns], f32)
    t = ns.astype(f32) * dt
    beta = (f0 - f0) * f32(1.0) / f32(release.get("fPulseW"))
    phase = (beta / f32(2) * (t * t) + f0 * t) + f32(90) / f32(360)
    c = np.cos(2 * 3.14 * phase.astype(np.float64)).astype(f32)
    expect = (env / f32(9)) * c
    assert np.array_equal(p.src_table[ns, 0], expect)


def test_cpml_tables_closed_form(release):
    """Roden-Gedney CPML: sigma graded (d/(P-1))^m, kappa 1 + (kmax-1)(.)^m,
    alpha graded linearly, b = exp(-(sigma/kappa + alpha) dt/eps0),
    c = sigma (b-1) / (sigma kappa + kappa^2 alpha); H values at half cells."""
    p = release.problem()
    P, dt, dx = p.npml, p.dt, p.dx
    m, kmax, amax = 3, 15.0, 0.24
    smax = 0.75 * 0.8 * (m + 1) / (dx * ETA0)
    i = np.arange(P)
    x = (P - 1 - i) / (P - 1.0)
    sig, kap, alp = smax * x ** m, 1 + (kmax - 1) * x ** m, amax * i / (P - 1.0)
    b = np.exp(-(sig / kap + alp) * dt / EPS0)
    c = sig * (b - 1) / (sig * kap + kap * kap * alp)
    assert np.allclose(p.be_x[0], b, rtol=1e-5, atol=0)
    assert np.allclose(p.ce_x[0], c, rtol=1e-4, atol=1e-9)
    ih = np.arange(P - 1)
    xh = (P - 1 - ih - 0.5) / (P - 1.0)
    sh, kh, ah = smax * xh ** m, 1 + (kmax - 1) * xh ** m, amax * (ih + 0.5) / (P - 1.0)
    bh = np.exp(-(sh / kh + ah) * dt / EPS0)
    ch = sh * (bh - 1) / (sh * kh + kh * kh * ah)
    assert np.allclose(p.bh_y[1], bh, rtol=1e-5)
    assert np.allclose(p.ch_y[1], ch, rtol=1e-4, atol=1e-9)
    # invariants: 0 < b <= 1, c <= 0, inner edge lossless, both sides equal
    for arr in (p.be_x, p.bh_x, p.be_y, p.bh_y):
        assert np.all((arr[0] > 0) & (arr[0] <= 1))
        assert np.array_equal(arr[0], arr[1])
    for arr in (p.ce_x, p.ch_x, p.ce_y, p.ch_y):
        assert np.all(arr[0] <= 0)
    assert p.ce_x[0][P - 1] == 0.0
    # den: 1/(kappa dx) inside the strips, mirrored on the high side, 1/dx inside
    assert p.den_ex[0] == pytest.approx(1 / (kap[0] * dx), rel=1e-6)
    assert p.den_ex[p.nx - 2] == pytest.approx(1 / (kap[1] * dx), rel=1e-6)
    assert p.den_hx[p.nx - 2] == pytest.approx(1 / (kh[0] * dx), rel=1e-6)
    assert np.all(p.den_ex[P:p.nx - P] == np.float32(1 / dx))


def test_material_table(release):
    p = release.problem()
    dt = p.dt
    assert p.pec_material == 1 and len(p.ca) == 150
    assert p.cb[0] == np.float32(dt / EPS0) and p.ca[0] == 1.0
    assert p.db == np.float32(dt / MU0)
    s, e = 5.76e7, 9.5 * EPS0  # copper
    assert p.ca[4] == np.float32((1 - s * dt / (2 * e)) / (1 + s * dt / (2 * e)))
    ne0 = 5.5e19
    assert p.ne[8] == 0 and p.ne[108] == pytest.approx(ne0, rel=1e-6)
    assert p.ne[50] == np.float32((42 * 0.01) * float(np.float32(ne0)))  # double product, float store


def test_medium_structure_and_cutoff_against_rc_line(release):
    """The plasma ellipse of the shipped Settings.ini: 2x2 Yee blocks, peak
    level 108 at the centre, and - the reference-held check - the 38 GHz
    cutoff (ne = nc) on the array axis where RC_Line.txt finds the edge."""
    p = release.problem()
    ids = p.material
    assert ids.max() == 108 and ids[2800, 250] == 108
    blocks = ids[:2998:2, :498:2]
    assert np.array_equal(blocks, ids[1:2999:2, :498:2]) and np.array_equal(blocks, ids[:2998:2, 1:499:2])
    ne_axis = p.ne[ids[:, 250]]
    nc = EPS0 * (2 * math.pi * 38e9) ** 2 / CJ
    j_cut = int(np.argmax(ne_axis >= nc))
    rc = json.loads((GOLD / "reference_lines.json").read_text())["rc_line"]
    edges = [r[0] for r in rc]
    assert min(edges) - 40 <= j_cut <= max(edges) + 5


def test_missing_ini_falls_back_to_compile_time_defaults(tmp_path):
    s = F.Setup.from_ini(str(tmp_path / "absent.ini"))
    assert s.ini_status == -1
    p = s.problem()
    assert (p.nx, p.ny, p.nsteps) == (550, 600, 10000)  # I_SIZE, J_SIZE, MAX_TIME_STEP with DS = 1


def test_plain_workload_parameters():
    """400x400, 1 mm, dt = 0.99 x 2D Courant, CPML 12 with
    sigma_max = 0.8(m+1)/(eta0 dx), kappa_max = 1, linear ramp to 2 nc(10 GHz)
    from x = 200 to 388, nu = 1e9, detectors (60, 170..230)."""
    s = F.Setup.workload(1, seed=1)
    p = s.problem()
    assert (p.nx, p.ny, p.npml, p.nsteps) == (400, 400, 12, 2000)
    assert p.dt == pytest.approx(0.99 / (C0 * math.sqrt(2) / 1e-3), rel=1e-12)
    nc = EPS0 * (2 * math.pi * 10e9) ** 2 / CJ
    assert s.get("nc") == pytest.approx(nc, rel=1e-7) and nc == pytest.approx(1.2405e18, rel=1e-3)
    assert p.density[7, 200] == 0 and p.density[7, 388] == np.float32(2 * nc)
    assert p.density[7, 294] == pytest.approx(nc, rel=1e-2)
    assert list(p.det_j) == [170, 190, 210, 230] and set(p.det_i) == {60}
    nu = 1e9
    assert p.jz_decay == np.float32((1 - nu * p.dt / 2) / (1 + nu * p.dt / 2))
    smax = 0.8 * 4 / (ETA0 * 1e-3)
    assert p.be_x[0][0] == pytest.approx(math.exp(-smax * p.dt / EPS0), rel=1e-6)
    # kappa_max = 1: den is 1/dx everywhere
    assert np.all(p.den_ex == np.float32(1e3))


def test_workload_shots_are_seeded_and_varied():
    a = F.Setup.workload(3, seed=5, shot=17).problem()
    b = F.Setup.workload(3, seed=5, shot=17).problem()
    c = F.Setup.workload(3, seed=5, shot=18).problem()
    assert (a.nx, a.ny, a.dt) == (b.nx, b.ny, b.dt) and np.array_equal(a.src_table, b.src_table)
    assert (a.nx, a.ny, len(a.src_i)) != (c.nx, c.ny, len(c.src_i)) or a.dt != c.dt
    for q in (a, c):
        assert q.nx in range(256, 2049, 128) and q.ny in range(256, 2049, 128)
        assert 1 <= len(q.src_i) <= 16 and q.nsteps == 4000


def test_gaussian_field_statistics():
    """The workloads' density fluctuation: unit variance, correlation
    exp(-r^2/l^2) with l = 20 cells (checked at lag 20 -> e^-1)."""
    p = F.Setup.workload(2, seed=2, steps=10).problem()
    d = p.density.astype(np.float64)
    prof = d.mean(axis=0)
    cols = np.arange(4500, 7500)
    g = (d[:, cols] / prof[cols] - 1) / 0.02
    assert g.std() == pytest.approx(1.0, rel=0.05)
    lag = 20
    r = np.mean(g[:, :-lag] * g[:, lag:]) / g.var()
    assert r == pytest.approx(math.exp(-1), abs=0.05)


def test_synthetic_shots_respect_the_yee_drude_stability_bound():
    """Shots of the batched and served workloads (dx = lambda/20, plasma up to
    3x critical) take dt under (c dt)^2 (1/dx^2 + 1/dy^2) + (wp dt/2)^2 <= 1
    for their densest cell; shot 176 of the batch mix (2.7x critical at
    22.7 GHz) diverged at 0.99 of the plain Courant limit."""
    import numpy as np

    c = 299792458.0
    for w, shot in [(3, 176), (3, 0), (3, 1), (4, 0), (5, 0)]:
        s = F.Setup.workload(w, seed=w, shot=shot, steps=10)
        p = s.c_problem.contents
        dens = np.ctypeslib.as_array(p.density, shape=(p.ny * p.nx,))
        wp2 = float(dens.max()) * 2.81794e-8 / EPS0
        bound = (c * p.dt) ** 2 * (1 / p.dx ** 2 + 1 / p.dy ** 2) + wp2 * p.dt ** 2 / 4
        assert bound <= 1.0, (w, shot, bound)
        s.close()


def test_rc_lines_follow_reference_positions(release):
    """calculateRCdetect2Darrai on the scaled [RC_calc] keys: Iline 125 mm,
    LIsize 100 mm, LIdelta 5 mm at ds 0.5 mm -> 21 lines from column 150 in
    steps of 10; the edge scan column adds the line index; the maximum search
    starts 150 rows past the source row; Rlevel 0.07; edge level from 10."""
    p = release.c_problem.contents
    assert p.nrc == 21
    imax = [p.rc_imax[e] for e in range(21)]
    isc = [p.rc_i[e] for e in range(21)]
    assert imax == [150 + 10 * e for e in range(21)]
    assert isc == [150 + 11 * e for e in range(21)]
    assert p.rc_jcut == int(release.get("Jsource")) + 150 == 250
    assert abs(p.rc_level - 0.07) < 1e-7 and p.rc_edge0 == 10.0


def _rc_post(spd, window, ds, dt):
    """RClineFilterPostTime, CalculateRCfromPwrLines, CalculateTaufromPwrLines
    (F/kernel.cu:3180-3270) restated with float32 sums."""
    nmax, q = spd.shape[0] - 1, spd.shape[1]
    weight = np.float32(1.0 / abs(window))
    fl = np.zeros((nmax, q), np.float32)
    for k in range(q):
        wb, we = window // 2, -(window // 2)
        for i in range(nmax):
            acc = np.float32(0)
            for j in range(wb, we, -1):
                if 0 < j < nmax:
                    acc = np.float32(acc + np.float32(spd[j, k]) * weight)
            fl[i, k] = acc
            wb += 1
            we += 1
    rc, tau = [], []
    for k in range(q):
        temp = max(0, int(spd[:nmax, k].max()))
        ft = max(np.float32(0), fl[:, k].max())
        rc.append((temp, float(ft), float(ft) * ds))
        i = 2000
        while True:
            f = float(np.float32(i * 2))
            i += 1
            if not (i < nmax and fl[i, k] - fl[i - 1, k] >= 0):
                break
        tau.append((f, f * dt * 1e12))
    return rc, tau


def test_rc_outputs_known_answers(release, tmp_path):
    """A constant line gives its own value raw and filtered, and the filtered
    line stops rising where the moving window starts to leave the run:
    TAU = 2 (nMax - window/2 - 1)."""
    p = release.c_problem.contents
    nmax, q = p.nsteps, p.nrc
    spd = np.zeros((nmax + 1, q), np.int32)
    spd[1:, :] = 100 + np.arange(q, dtype=np.int32)[None, :]
    release.write_rc_outputs(str(tmp_path), spd)
    rc = [l.split() for l in (tmp_path / "RC_Line.txt").read_text().splitlines()]
    tau = [l.split() for l in (tmp_path / "TAU_Line.txt").read_text().splitlines()]
    assert len(rc) == len(tau) == q
    w = int(release.get("FilterWindow"))
    for k in range(q):
        assert int(rc[k][0]) == 100 + k
        assert abs(float(rc[k][1]) - (100 + k)) < 1e-3 * (100 + k)
        assert float(tau[k][0]) == 2 * (nmax - w // 2 - 1)
    pos = (tmp_path / "RcDet_data.txt").read_text().split("\n")
    assert [int(x) for x in pos[0].split()] == [150 + 10 * e for e in range(21)]


def test_rc_outputs_match_restatement(release, tmp_path):
    """A moving, noisy line through the writers against the restatement above
    (formatting as std::ostream prints float and double)."""
    p = release.c_problem.contents
    nmax, q = p.nsteps, p.nrc
    rng = np.random.default_rng(5)
    t = np.arange(nmax + 1)[:, None]
    spd = (2000 + 0.05 * t + 20 * np.sin(t / 700.0 + np.arange(q)[None, :]) + rng.integers(-3, 4, (nmax + 1, q)))
    spd = spd.astype(np.int32)
    spd[0] = 0
    release.write_rc_outputs(str(tmp_path), spd)
    rc_want, tau_want = _rc_post(spd[:, 4:5], int(release.get("FilterWindow")), release.get("ds"), p.dt)
    rc = [l.split() for l in (tmp_path / "RC_Line.txt").read_text().splitlines()][4:5]
    tau = [l.split() for l in (tmp_path / "TAU_Line.txt").read_text().splitlines()][4:5]
    for k in range(1):
        assert int(rc[k][0]) == rc_want[k][0]
        assert rc[k][1] == f"{rc_want[k][1]:.6g}" and rc[k][2] == f"{rc_want[k][2]:.6g}"
        assert tau[k][0] == f"{tau_want[k][0]:.6g}" and tau[k][1] == f"{tau_want[k][1]:.6g}"
