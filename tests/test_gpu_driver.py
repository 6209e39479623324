"""The Fortran host driver (parameter.d in, reference restart files out) on the GPU:
10 steps in one run equal 5 steps, a restart through vel2D<idp>.0000 / part.0000
(load_save.f90 formats), and 5 more steps; a restart from the per-rank files of a
2 x 2 decomposition equals the restart from the one-domain files; the statistics and
collision files in the reference's formats."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
DRIVER = ROOT / "mpi_2d_domain_decomp_turb_particles_amd" / "turb_gpu"
BASE = ROOT / "tests" / "golden" / "parameter.d_64_short"


def _param(path, nstep, new_flow, idp, new_particle, nprocY=None):
    lines = BASE.read_text().split("\n")
    if nprocY is not None:
        lines[2] = str(nprocY)
    lines[4] = str(nstep)
    lines[12] = ".TRUE." if new_flow else ".FALSE."
    lines[14] = str(idp)
    lines[44] = ".TRUE." if new_particle else ".FALSE."
    lines[48] = ".FALSE."          # no HI: the restart files carry no disturbance velocity
    lines[52] = "5000"
    path.write_text("\n".join(lines))


def _read_records(path, count):
    """Fortran unformatted sequential records (4-byte length markers)."""
    out = []
    with open(path, "rb") as f:
        for _ in range(count):
            n = int(np.frombuffer(f.read(4), dtype=np.int32)[0])
            out.append(f.read(n))
            assert int(np.frombuffer(f.read(4), dtype=np.int32)[0]) == n
    return out


def _run(cwd, param, *args):
    r = subprocess.run([str(DRIVER), str(param), *args], cwd=cwd, capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0 and "Successfully completed" in r.stdout


@pytest.mark.skipif(not DRIVER.exists(), reason="Fortran driver not built")
def test_restart_through_reference_files(tmp_path):
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir(); b.mkdir()
    _param(a / "p10", 10, True, 0, True)
    _run(a, a / "p10")
    _param(b / "p5", 5, True, 0, True)
    _run(b, b / "p5")
    _param(b / "p5r", 5, False, 1, False)
    _run(b, b / "p5r")
    va = _read_records(a / "vel2D1.0000", 6)
    vb = _read_records(b / "vel2D2.0000", 6)
    # a restart rebuilds the z-inverse fields from the stored (symmetrised) u-hat, the
    # continuous run symmetrises the kx = 0 plane inside the y stage: the same
    # operations in another order, so the fields agree to rounding
    for x, y in zip(va, vb):
        fx, fy = np.frombuffer(x, dtype=np.float64), np.frombuffer(y, dtype=np.float64)
        assert np.abs(fx - fy).max() <= 1e-12 * np.abs(fx).max()
    pa = _read_records(a / "part.0000", 2)
    pb = _read_records(b / "part.0000", 2)
    assert pa[0] == pb[0]                               # same droplet count
    n = int(np.frombuffer(pa[0], dtype=np.int32)[0])
    ya = np.frombuffer(pa[1], dtype=np.float64)[: 8 * n].reshape(n, 8)
    yb = np.frombuffer(pb[1], dtype=np.float64)[: 8 * n].reshape(n, 8)
    assert np.array_equal(ya[:, 7], yb[:, 7])
    assert np.abs(ya - yb).max() < 1e-10


@pytest.mark.skipif(not DRIVER.exists(), reason="Fortran driver not built")
def test_statistics_files(tmp_path):
    """The driver writes POSTPROC's monitora/monitorb/monitorc.dat and the spectrum and
    transfer files of TRANSPOSTPROC in the reference's layout (postprocessing.f90:157-315):
    8, 8 and 6 columns, and a shell table per output; monitorb's total energy and
    dissipation are the sums of the spectrum.dat columns of the same step, and the
    transfer sums to zero (energy conservation of the dealiased nonlinear term)."""
    lines = BASE.read_text().split("\n")
    lines[4] = "4"             # nstep
    lines[24] = "2"            # ieout
    lines[26] = "2"            # ispec
    lines[28] = ".TRUE."       # divstat
    (tmp_path / "p").write_text("\n".join(lines))
    _run(tmp_path, tmp_path / "p")
    ma = np.loadtxt(tmp_path / "monitora.dat", ndmin=2)
    mb = np.loadtxt(tmp_path / "monitorb.dat", ndmin=2)
    mc = np.loadtxt(tmp_path / "monitorc.dat", ndmin=2)
    assert ma.shape == (2, 8) and mb.shape == (2, 8) and mc.shape == (2, 6)
    assert np.all(np.isfinite(ma)) and np.all(np.isfinite(mb)) and np.all(np.isfinite(mc))
    assert np.all(mb[:, 7] > 1.0)                         # flatness of a Gaussian-like field
    nek = int(64 / 2 - 1.5)                              # turb.f90:303-304
    spec = (tmp_path / "spectrum.dat").read_text().split()
    tran = (tmp_path / "transfer.dat").read_text().split()
    for out in range(2):
        s = np.array(spec[out * (1 + 3 * nek) + 1: (out + 1) * (1 + 3 * nek)], dtype=float).reshape(nek, 3)
        t = np.array(tran[out * (1 + 2 * nek) + 1: (out + 1) * (1 + 2 * nek)], dtype=float).reshape(nek, 2)
        assert np.array_equal(s[:, 0], np.arange(1, nek + 1))
        assert abs(s[:, 1].sum() - mb[out, 1]) <= 1e-4 * mb[out, 1]      # list-directed and 1pe15.5 rounding
        assert abs(s[:, 2].sum() - mb[out, 2]) <= 1e-4 * mb[out, 2]
        assert abs(t[:, 1].sum()) <= 1e-9 * np.abs(t[:, 1]).sum()


@pytest.mark.skipif(not DRIVER.exists(), reason="Fortran driver not built")
def test_collision_statistics_files(tmp_path):
    """COLLISION_DET's files in the reference's formats (partlib.f90:1641-1797): the
    summed bins every fifth step, 180 per line, thetabin a mean angle in [0, pi/2];
    collision.dat lines whose type counts and size-class counts add up to the same
    total; one Vp<i1><i2>.dat line per contact that detectpairs.dat counts, carrying
    a unit offset and the relative velocity along it."""
    lines = BASE.read_text().split("\n")
    lines[4] = "9"             # nstep: bins written after steps 4 and 9
    lines[48] = ".FALSE."      # no HI
    lines[52] = "100000"
    (tmp_path / "p").write_text("\n".join(lines))
    _run(tmp_path, tmp_path / "p")
    total = 0
    for pair in ("11", "12", "22"):
        nir = np.loadtxt(tmp_path / f"nirbin{pair}.dat", ndmin=2)
        assert nir.shape == (2, 180) and np.all(nir >= 0)
        total += nir.sum()
        for name in ("wrbin", "wxbin", "wybin", "wzbin", "urbin", "thetabin", "wrtanbin"):
            b = np.loadtxt(tmp_path / f"{name}{pair}.dat", ndmin=2)
            assert b.shape == (2, 180) and np.all(np.isfinite(b)), name
        th = np.loadtxt(tmp_path / f"thetabin{pair}.dat", ndmin=2)
        assert np.all(th >= 0) and np.all(th <= np.pi / 2 + 1e-4)
        assert np.all(th[nir == 0] == 0)
    assert total > 1000
    coll = np.loadtxt(tmp_path / "collision.dat", ndmin=2)
    if coll.size:
        assert coll.shape[1] == 7
        assert np.array_equal(coll[:, 1:4].sum(1), coll[:, 4:7].sum(1))
    det = np.loadtxt(tmp_path / "detectpairs.dat", ndmin=2)
    assert det.shape[1] == 4 and np.all(det[:, 1:].sum(1) > 0)
    for k, pair in enumerate(("11", "12", "22")):
        vp = np.loadtxt(tmp_path / f"Vp{pair}.dat", ndmin=2)
        assert len(vp) == det[:, 1 + k].sum()
        if len(vp):
            assert vp.shape[1] == 11
            assert np.abs(np.linalg.norm(vp[:, 7:10], axis=1) - 1).max() < 5e-5
            wr = ((vp[:, 1:4] - vp[:, 4:7]) * vp[:, 7:10]).sum(1)
            assert np.abs(wr - vp[:, 10]).max() < 5e-5 * (1 + np.abs(vp[:, 1:7]).max())
    assert det[:, 1:].sum() > 0


def _write_records(path, recs):
    """Fortran unformatted sequential records (4-byte length markers)."""
    with open(path, "wb") as f:
        for r in recs:
            m = np.array([len(r)], dtype=np.int32).tobytes()
            f.write(m)
            f.write(r)
            f.write(m)


def split_velocity_file(src, dst_dir, idump, N, npy, npz):
    """The per-rank files vel2D<idump>.<rank> that an nprocY x nprocZ run of the
    reference writes (load_save.f90:2-55), made from a one-domain file src (six records
    v(N+2, N+2, N)): rank id = iy + nprocY iz holds v(:, iy ly + 1 : iy ly + ly + lyext,
    iz lz + 1 : iz lz + lz), ly = N / nprocY, lz = N / nprocZ, with the two extra rows
    (lyext = 2) on the last rank along y (turb.f90:282-295)."""
    recs = _read_records(src, 6)
    L = N + 2
    ly, lz = N // npy, N // npz
    for rank in range(npy * npz):
        iy, iz = rank % npy, rank // npy
        lyx = 2 if iy == npy - 1 else 0
        out = []
        for r in recs:
            v = np.frombuffer(r, dtype=np.float64).reshape((L, L, N), order="F")
            blk = v[:, iy * ly: iy * ly + ly + lyx, iz * lz: iz * lz + lz]
            out.append(blk.tobytes(order="F"))
        _write_records(dst_dir / f"vel2D{idump}.{rank:04d}", out)


def split_droplet_file(src, dst_dir, nranks):
    """part.<rank> files (load_save.f90:99-124: the count, then yp(8, n), fvp(3, 0:2, n)
    and fdp(3, 0:3, n) in one record) holding consecutive slices of the droplets of
    the one-domain file src, so that reading them in rank order gives the same list."""
    head, body = _read_records(src, 2)
    n = int(np.frombuffer(head, dtype=np.int32)[0])
    x = np.frombuffer(body, dtype=np.float64)
    yp = x[: 8 * n].reshape((8, n), order="F")
    fvp = x[8 * n: 17 * n].reshape((9, n), order="F")
    fdp = x[17 * n: 29 * n].reshape((12, n), order="F")
    cuts = np.linspace(0, n, nranks + 1).astype(int)
    for rank in range(nranks):
        a, b = cuts[rank], cuts[rank + 1]
        rec = b"".join(np.asfortranarray(arr[:, a:b]).tobytes(order="F") for arr in (yp, fvp, fdp))
        _write_records(dst_dir / f"part.{rank:04d}", [np.array([b - a], dtype=np.int32).tobytes(), rec])


@pytest.mark.skipif(not DRIVER.exists(), reason="Fortran driver not built")
def test_restart_from_four_rank_files(tmp_path):
    """A restart from the files of a 2 x 2 decomposition (nprocY = nprocZ = 2: four
    vel2D1.<rank> and part.<rank> files) resumes to exactly the state of the restart
    from the one-domain files they were cut from."""
    import shutil
    a, b, c = tmp_path / "a", tmp_path / "b", tmp_path / "c"
    for d in (a, b, c):
        d.mkdir()
    _param(a / "p5", 5, True, 0, True)
    _run(a, a / "p5")
    shutil.copy(a / "vel2D1.0000", b / "vel2D1.0000")
    shutil.copy(a / "part.0000", b / "part.0000")
    _param(b / "p5r", 5, False, 1, False)
    _run(b, b / "p5r")
    N = 64
    split_velocity_file(a / "vel2D1.0000", c, 1, N, 2, 2)
    split_droplet_file(a / "part.0000", c, 4)
    assert sorted(p.name for p in c.iterdir()) == [f"part.{r:04d}" for r in range(4)] + \
        [f"vel2D1.{r:04d}" for r in range(4)]
    _param(c / "p5r", 5, False, 1, False, nprocY=2)
    _run(c, c / "p5r", "fp64", "4")
    for name, count in (("vel2D2.0000", 6), ("part.0000", 2)):
        rb, rc = _read_records(b / name, count), _read_records(c / name, count)
        assert all(x == y for x, y in zip(rb, rc)), name
