"""The C++ host side: the reference-shaped adapter (cfdv_amd/host/
cfdv_reference_adapter.hpp) and the deck driver cfdv_run built on it.

CPU: the driver's deck loader + the adapter's flattening hand cfdv_create
exactly the arrays the Python deck reader produces (connectivity, face bc
codes, coordinates, Dirichlet flags of add_face, conservative U).
GPU: two reference-mode iterations of the driver on 3dtestramp_new match
the oracle's golden states, and RMSError.txt has MPI_RMSError's format.
"""
import os
import subprocess

import numpy as np
import pytest

from cfdv_amd import build_lib
from cfdv_amd import mesh as M

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DECKS = ["3dtestlagrange", "3dtestramp_new", "lid50"]


@pytest.fixture(scope="module")
def runner():
    if not os.path.exists(build_lib.RUNNER):
        if not os.path.exists(build_lib.LIB):
            build_lib.build()
        else:
            build_lib.build_runner()
    return build_lib.RUNNER


@pytest.mark.parametrize("name", DECKS)
def test_driver_flattens_deck_like_the_reader(runner, tmp_path, name):
    deck = os.path.join(GOLD, "decks", name)
    out = tmp_path / "flat.bin"
    subprocess.run([runner, deck, "--flatten", str(out)], check=True, capture_output=True)
    mesh, th, U, _ = M.read_deck(deck)
    raw = out.read_bytes()
    ne, nn, nnod, nf, nd, nq = mesh.n_elem, mesh.n_node, mesh.nnod, mesh.nface, mesh.ndim, mesh.neqn
    o = 0
    conn = np.frombuffer(raw, np.int32, ne * nnod, o).reshape(ne, nnod)
    o += 4 * ne * nnod
    bc = np.frombuffer(raw, np.int32, ne * nf, o).reshape(ne, nf)
    o += 4 * ne * nf
    x = np.frombuffer(raw, np.float64, nn * nd, o).reshape(nn, nd)
    o += 8 * nn * nd
    d = np.frombuffer(raw, np.uint8, nn * nq, o).reshape(nn, nq)
    o += nn * nq
    Uc = np.frombuffer(raw, np.float64, nn * nq, o).reshape(nn, nq)
    assert o + 8 * nn * nq == len(raw)
    assert np.array_equal(conn, mesh.conn)
    assert np.array_equal(bc, mesh.bc)
    assert np.array_equal(x, mesh.x)
    assert np.array_equal(d, mesh.dirichlet)
    assert np.array_equal(Uc, U)


def test_driver_reports_missing_device_or_deck(runner, tmp_path):
    r = subprocess.run([runner, str(tmp_path / "nope")], capture_output=True, text=True)
    assert r.returncode == 1 and "cannot open" in r.stderr


@pytest.mark.gpu
def test_driver_two_reference_iterations(gpu_available, runner, tmp_path):
    name = "3dtestramp_new"
    g = np.load(os.path.join(GOLD, f"{name}.npz"))
    dump = tmp_path / "U.bin"
    r = subprocess.run([runner, os.path.join(GOLD, "decks", name), "--iters", "2",
                        "--out", str(tmp_path), "--dump", str(dump)],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    U2 = np.fromfile(dump, dtype=np.float64).reshape(g["U0"].shape)
    ref = g["U_steps"][1]
    err = np.linalg.norm(U2 - ref) / np.linalg.norm(ref - g["U0"])
    print("U after 2 iterations, error relative to the change:", err)
    assert err < 1e-9
    lines = (tmp_path / "RMSError.txt").read_text().splitlines()
    assert len(lines) == 2
    for k, line in enumerate(lines):
        it, vals = line.split("\t")
        assert int(it) == k
        v = np.array([float(t) for t in vals.split()])
        assert np.allclose(v, g["rms_steps"][k], rtol=1e-5, atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_driver_solvers_on_rows_and_elements(gpu_available, runner, tmp_path, mode):
    """One iteration of the FP64 and mixed solvers on the ramp deck, products
    over node rows (the default) and element by element: both reach the
    golden textbook solution of the first step's system (du_true, a dense
    solve of the oracle's K) and agree with each other."""
    name = "3dtestramp_new"
    g = np.load(os.path.join(GOLD, f"{name}.npz"))
    out = {}
    for storage in ("rows", "ebe"):
        d = tmp_path / storage
        d.mkdir()
        r = subprocess.run([runner, os.path.join(GOLD, "decks", name), "--iters", "1", "--mode", mode,
                            "--rtol", "1e-10", "--precond", "nodeblock", "--storage", storage,
                            "--out", str(d), "--dump", str(d / "U.bin")],
                           capture_output=True, text=True, timeout=300)
        print(storage, r.stdout[-400:], r.stderr[-400:])
        assert r.returncode == 0
        out[storage] = np.fromfile(d / "U.bin", dtype=np.float64).reshape(g["U0"].shape) - g["U0"]
    ref = g["du_true"].reshape(g["U0"].shape)
    for storage, du in out.items():
        err = np.linalg.norm(du - ref) / np.linalg.norm(ref)
        print(mode, storage, "du vs textbook solution", err)
        assert err < 1e-7
    assert np.linalg.norm(out["rows"] - out["ebe"]) <= 1e-7 * np.linalg.norm(ref)


@pytest.mark.gpu
def test_driver_warm_start(gpu_available, runner, tmp_path):
    """--warm-start 1 and 2 (fp64/mixed modes) change where each solve starts,
    not where it stops: four mixed iterations on the ramp deck land on the
    cold-started iterations' state to the solver tolerance.  Other values
    are refused."""
    name = "3dtestramp_new"
    deck = os.path.join(GOLD, "decks", name)
    U = {}
    for ws in (0, 1, 2):
        d = tmp_path / f"ws{ws}"
        d.mkdir()
        r = subprocess.run([runner, deck, "--iters", "4", "--mode", "mixed", "--rtol", "1e-10", "--precond",
                            "nodeblock", "--warm-start", str(ws), "--out", str(d), "--dump", str(d / "U.bin")],
                           capture_output=True, text=True, timeout=300)
        print(ws, r.stdout[-300:], r.stderr[-300:])
        assert r.returncode == 0
        U[ws] = np.fromfile(d / "U.bin", dtype=np.float64)
    g = np.load(os.path.join(GOLD, f"{name}.npz"))
    dU = np.linalg.norm(U[0] - g["U0"].ravel())
    for ws in (1, 2):
        err = np.linalg.norm(U[ws] - U[0]) / dU
        print("warm start", ws, "vs cold, relative to the 4-step change:", err)
        assert 0 < err <= 1e-7      # a different start, the same stopping test
    r = subprocess.run([runner, deck, "--iters", "1", "--mode", "mixed", "--warm-start", "3",
                        "--out", str(tmp_path)], capture_output=True, text=True, timeout=300)
    assert r.returncode != 0 and "warm-start" in (r.stdout + r.stderr)


@pytest.mark.gpu
def test_driver_checkpoint_restart_and_time_file(gpu_available, runner, tmp_path):
    """--save 1 writes a restartable deck before every iteration and after
    the last one (config->Save, main.cpp:177, 208); running the driver on the
    iteration-1 checkpoint for one iteration lands where two uninterrupted
    iterations land (the saved primitives round-trip U to rounding); time.txt
    has IterationTimer's header and one line per iteration."""
    deck = os.path.join(GOLD, "decks", "3dtestlagrange")
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    r = subprocess.run([runner, deck, "--iters", "2", "--save", "1", "--out", str(a), "--dump", str(a / "U.bin")],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout[-500:], r.stderr[-500:])
    assert r.returncode == 0
    for it in (0, 1, 2):
        assert (a / str(it) / "binarynodes").exists() and (a / str(it) / "binaryelements").exists()
        ctl = (a / str(it) / "control").read_text()
        assert f"iter = {it}" in ctl
    m0, _, U0, _ = M.read_deck(deck)
    m2, _, U2d, _ = M.read_deck(str(a / "2"))
    assert np.array_equal(m2.conn, m0.conn) and np.array_equal(m2.bc, m0.bc)
    UA = np.fromfile(a / "U.bin", dtype=np.float64).reshape(U0.shape)
    assert np.allclose(U2d, UA, rtol=1e-13, atol=1e-15)
    r = subprocess.run([runner, str(a / "1"), "--iters", "1", "--save", "0", "--out", str(b),
                        "--dump", str(b / "U.bin")], capture_output=True, text=True, timeout=300)
    print(r.stdout[-500:], r.stderr[-500:])
    assert r.returncode == 0
    UB = np.fromfile(b / "U.bin", dtype=np.float64).reshape(U0.shape)
    _, _, U1d, _ = M.read_deck(str(a / "1"))
    err = np.linalg.norm(UB - UA) / np.linalg.norm(UA - U1d)
    print("restart vs uninterrupted, relative to the step's change:", err)
    assert err < 1e-9
    assert (b / "RMSError.txt").read_text().split("\t")[0] == "1"
    lines = (a / "time.txt").read_text().splitlines()
    assert lines[0] == "Iteration, Run_Time, Time_Elapsed, Sim_Time, Step_dt"
    assert len(lines) == 3
    rows = [[float(v) for v in ln.split()] for ln in lines[1:]]
    assert [int(r_[0]) for r_ in rows] == [0, 1]
    assert all(len(r_) == 5 and r_[1] >= r_[2] > 0 and r_[4] > 0 for r_ in rows)
    # six significant digits, the stream default IterationTimer writes with
    assert rows[1][3] == pytest.approx(rows[0][3] + rows[1][4], rel=1e-5)


def _refined_deck(tmp_path):
    """a flat-plate deck refined twice on the host (3D node matching)"""
    from cfdv_amd import refine as PR
    mesh, th, U = M.flat_plate(6, 3, 3)
    r = PR.RefinableMesh(mesh, U)
    for every, off in ((5, 2), (7, 1)):
        ne = r.sizes()[0]
        r.apply(*r.refine_list((np.arange(ne) % every == off).astype(np.uint8)), PR.MODE_3D)
    m2, U2, _, lv = r.get()
    m2.refine_level = lv
    p = M.primitives(U2, 3, th.Cv, th.gamma)
    deck = tmp_path / "refined"
    M.write_deck(str(deck), m2, th, p["rho"], p["v"], p["T"])
    return deck, m2, U2


def test_adap_deck_round_trip(runner, tmp_path):
    deck, m2, U2 = _refined_deck(tmp_path)
    back, th, U, ctl = M.read_deck(str(deck))
    assert int(ctl["adap"]) == 1
    assert np.array_equal(back.conn, m2.conn) and np.array_equal(back.bc, m2.bc)
    assert np.array_equal(back.adap, m2.adap) and np.array_equal(back.refine_level * back.adap,
                                                                  m2.refine_level * m2.adap)
    sel = m2.adap.astype(bool)
    assert np.array_equal(back.hnm[sel], m2.hnm[sel])
    assert np.allclose(U, U2, rtol=1e-14, atol=0)
    # the C++ driver reads the same adap file into Element::adap / Hnm
    out = tmp_path / "flat.bin"
    subprocess.run([runner, str(deck), "--flatten", str(out)], check=True, capture_output=True)
    raw = out.read_bytes()
    conn = np.frombuffer(raw, np.int32, back.n_elem * 8, 0).reshape(-1, 8)
    assert np.array_equal(conn, back.conn)


@pytest.mark.gpu
def test_driver_adap_pass_matches_host_refinement(gpu_available, runner, tmp_path):
    """cfdv_run --adap 3d (device flags, host closure and split, SaveBinary)
    writes the deck that the Python mirror of the same pass produces"""
    from cfdv_amd import refine as PR
    from cfdv_amd.fdv import FDVContext
    mesh, th, U = M.flat_plate(8, 4, 4)
    p = M.primitives(U, 3, th.Cv, th.gamma)
    src = tmp_path / "plate"
    M.write_deck(str(src), mesh, th, p["rho"], p["v"], p["T"])
    mesh, th, U, _ = M.read_deck(str(src))
    out = tmp_path / "plate_refined"
    os.makedirs(out)
    r = subprocess.run([runner, str(src), "--adap", "3d", "--out", str(out)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    ctx = FDVContext(mesh, th, U)
    flags, _, _ = ctx.refine_flags()
    assert flags.any()
    m2, U2, _, lv, el, rl = PR.refine_pass(mesh, U, flags, mode=PR.MODE_3D)
    got, _, Ug, _ = M.read_deck(str(out))
    assert np.array_equal(got.conn, m2.conn) and np.array_equal(got.bc, m2.bc)
    assert np.array_equal(got.x, m2.x)
    assert np.array_equal(got.adap, m2.adap)
    sel = m2.adap.astype(bool)
    assert np.array_equal(got.hnm[sel], m2.hnm[sel])
    assert np.array_equal(got.refine_level[sel], lv[sel])
    assert np.allclose(Ug, U2, rtol=1e-13, atol=0)
    # the refined deck runs: one reference-mode iteration of the driver
    r = subprocess.run([runner, str(out), "--iters", "1", "--out", str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr


def _unpack_deck(tmp_path, name):
    """a golden deck stored with gzipped binary records, unpacked to tmp"""
    import gzip
    import shutil
    src = os.path.join(GOLD, "decks", name)
    dst = tmp_path / name
    os.makedirs(dst)
    for f in os.listdir(src):
        if f.endswith(".gz"):
            with gzip.open(os.path.join(src, f), "rb") as fi, open(dst / f[:-3], "wb") as fo:
                fo.write(fi.read())
        else:
            shutil.copyfile(os.path.join(src, f), dst / f)
    return dst


def test_tecplot_writer_reproduces_reference_output(runner, tmp_path):
    """MPI_TecplotOut's file for backstep_refined_new at t = 0, byte for
    byte against the tecplot_0.plt the reference ships with that deck"""
    import gzip
    deck = _unpack_deck(tmp_path, "backstep_refined_new")
    out = tmp_path / "out"
    os.makedirs(out)
    subprocess.run([runner, str(deck), "--tecplot-only", "--out", str(out)], check=True, capture_output=True)
    got = (out / "tecplot_0.plt").read_bytes()
    with gzip.open(os.path.join(GOLD, "tecplot_backstep_0.plt.gz"), "rb") as f:
        want = f.read()
    assert len(got) == len(want)
    assert got == want


@pytest.mark.gpu
def test_tecplot_during_time_loop_has_element_s(gpu_available, runner, tmp_path):
    """with the time loop, the Output file also carries s1..s4 per element
    (ele_data, main.cpp:183): the node blocks equal the t = 0 file and the s
    blocks the oracle's FDVParam to the printed 7 digits"""
    from oracle import oracle as O
    deck = os.path.join(GOLD, "decks", "lid50")
    out = tmp_path / "out"
    os.makedirs(out)
    r = subprocess.run([runner, deck, "--iters", "1", "--tecplot", "1", "--out", str(out)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    only = tmp_path / "only"
    os.makedirs(only)
    subprocess.run([runner, deck, "--tecplot-only", "--out", str(only)], check=True, capture_output=True)
    a = (out / "tecplot_0.plt").read_text().split("\n")
    b = (only / "tecplot_0.plt").read_text().split("\n")
    assert a[1].startswith(b[1]) and a[1].endswith('"s1", "s2", "s3", "s4", ')
    assert a[3] == "VARLOCATION=([8,9,10,11]=CellCentered)"
    mesh, th, U, _ = M.read_deck(deck)
    # node blocks (x y rho u v T p): identical text
    body_a = "\n".join(a[4:]).split()
    body_b = "\n".join(b[3:]).split()
    nvals = 7 * mesh.n_node
    assert body_a[:nvals] == body_b[:nvals]
    s_file = np.array([float(t) for t in body_a[nvals:nvals + 4 * mesh.n_elem]]).reshape(4, -1).T
    so, _ = O.fdvparam_all(th, mesh, U)
    assert np.allclose(s_file, so, rtol=1e-6, atol=1e-300)
    conn = np.array([int(t) for t in body_a[nvals + 4 * mesh.n_elem:]]).reshape(-1, 4) - 1
    assert np.array_equal(conn, mesh.conn)
