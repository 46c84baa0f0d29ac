"""The Fortran host program (host/lbcuda_main.f90) keeps the reference's
input.dat / input.xyz semantics (readParameters main.CUF:409-496,
read_input_atom :1230-1350) and drives the C ABI.  On the CPU it is linked
against the oracle-backed implementation of the same ABI
(oracle/liblbm_cpu_abi.so); on a GPU, against liblbcuda_amd.so."""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def build_host(tmp, libdir, libname):
    out = os.path.join(tmp, "host_exe")
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "host"), f"OUT={out}",
                           f"LIBDIR={libdir}", f"LIBNAME={libname}", f"MODDIR={tmp}/mod"])
    return out


@pytest.fixture(scope="module")
def cpu_host(tmp_path_factory):
    if not os.path.exists("/opt/rocm/bin/amdflang"):
        pytest.skip("amdflang not available")
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    tmp = str(tmp_path_factory.mktemp("host"))
    return build_host(tmp, os.path.join(ROOT, "oracle"), "lbm_cpu_abi")


def write_case(d, input_dat, xyz=None, nml=None):
    os.makedirs(d, exist_ok=True)
    open(os.path.join(d, "input.dat"), "w").write(input_dat)
    if xyz:
        open(os.path.join(d, "input.xyz"), "w").write(xyz)
    if nml:
        open(os.path.join(d, "lbcuda.nml"), "w").write(nml)


def run(exe, d):
    return subprocess.run([exe], cwd=d, capture_output=True, text=True, timeout=300)


DROPLET = """20\t!niter  number of iterations
10
10\t\t!vtk print every n
1               ! init TYPE 1=droplet
0.0 0.0 0.0\t! vx,vy,vz
NO 0\t\t! with PARTICLES
0.0 0.0 0.0\t! ext_fxx,ext_fyy,ext_fzz
NO\t\t! with ROTATION
0.0 0.0 0.0\t! ext_tqx,ext_tqy,ext_tqz
"""


def test_droplet_deck_runs_and_conserves_mass(cpu_host, tmp_path):
    d = str(tmp_path / "drop")
    write_case(d, DROPLET, nml="&lbcuda box=16, sigma=0.05, beta=0.7 /\n")
    r = run(cpu_host, d)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("Iter:") == 2
    m0 = [float(x) for x in r.stdout.split("initial mass R,B:")[1].split()[:2]]
    m1 = [float(x) for x in r.stdout.split("final mass R,B:")[1].split()[:2]]
    assert np.allclose(m0, m1, rtol=1e-10)


def particle_deck(max_vel, n=4):
    return f"""12\t           \t!niter
6    \t\t       \t!stat print on screen every n
100\t\t\t!vtk print every n
6\t\t\t! init TYPE 6=RandomR+ b=1-R
0.0 0.0 0.0\t\t! vx,vy,vz
YES {n}\t\t\t! with PARTICLES
0.0 0.0 0.0\t\t! ext_fxx,ext_fyy,ext_fzz
MAX_PARTICLE_VEL {max_vel}\t! Max particle speed
YES\t\t\t! with ROTATION
0.0 0.0 0.0\t\t! ext_tqx,ext_tqy,ext_tqz
"""


XYZ = """     4
 read list vx vy vz phi psi theta
C   8.00  8.00  8.00  0.02  0.00  0.00   1.95724226  0.88751093 -2.51221137
C  24.00  8.00 24.00  0.00  0.02  0.00  -2.39282916 -0.95854742  0.56342170
C   8.00 24.00 24.00  0.00  0.00  0.02   1.60187667 -1.00965411  2.61776966
C  24.00 24.00  8.00  0.01  0.01  0.00   0.5         0.5         0.5
"""


def test_particle_deck_runs(cpu_host, tmp_path):
    d = str(tmp_path / "part")
    write_case(d, particle_deck(1.75), XYZ, "&lbcuda box=32, sigma=0.02, beta=0.7, wetting='neutral' /\n")
    r = run(cpu_host, d)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "particle 1 position:" in r.stdout
    x = [float(v) for v in r.stdout.split("particle 1 position:")[1].split()[:3]]
    # moves forward but less than free flight: the fluid at rest drags it
    assert 8.0 < x[0] < 8.0 + 12 * 0.02


def test_extreme_velocity_stops_the_run(cpu_host, tmp_path):
    """MAX_PARTICLE_VEL below the initial speed trips nve_lf
    (kernels_fluid_PART.CUF:1544-1548); the host stops with the error."""
    d = str(tmp_path / "vel")
    write_case(d, particle_deck(0.001), XYZ, "&lbcuda box=32, wetting='neutral' /\n")
    r = run(cpu_host, d)
    assert r.returncode != 0
    assert "lbm_step error    6" in r.stdout and "1545" in r.stdout


def test_incomplete_input_dat_is_rejected(cpu_host, tmp_path):
    """readParameters: a deck without the particle line fails (main.CUF:478-492)."""
    d = str(tmp_path / "bad")
    write_case(d, "10\n1\n1\n1\n0 0 0\n")
    r = run(cpu_host, d)
    assert r.returncode != 0
    assert "withParticles not found" in r.stdout


# ---- output and restart files (OutputVTK main.CUF:1812-1929, initType 4) ----

def read_vti(path):
    """Header text, Float32 payload and trailer of an appended-raw ImageData
    file (writeImageDataVTI, write_output.f90:115-391)."""
    raw = open(path, "rb").read()
    tag = b'<AppendedData encoding="raw">'
    i = raw.index(tag) + len(tag)
    j = raw.index(b"_", i) + 1
    nbytes = int(np.frombuffer(raw[j:j + 4], "<i4")[0])
    data = np.frombuffer(raw[j + 4:j + 4 + nbytes], "<f4")
    return raw[:i].decode(), data, raw[j + 4 + nbytes:].decode()


def read_init_bin(path, shape):
    """Unformatted sequential file of initType 3/4 (main.CUF:2444-2529):
    a record of six extents, then a record of the array (4-byte markers)."""
    raw = open(path, "rb").read()
    n0 = int(np.frombuffer(raw[0:4], "<i4")[0])
    ext = np.frombuffer(raw[4:4 + n0], "<i4")
    o = 4 + n0 + 4
    n1 = int(np.frombuffer(raw[o:o + 4], "<i4")[0])
    a = np.frombuffer(raw[o + 4:o + 4 + n1], "<f4").reshape(shape, order="F")
    assert int(np.frombuffer(raw[o + 4 + n1:o + 8 + n1], "<i4")[0]) == n1
    return ext, a


# D3Q19 velocities (dimensions_m.CUF:212-219)
E = np.array([[0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 0, 0, 0, 0],
              [0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1],
              [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1]], dtype=np.float64).T


def oracle_moments(o):
    """moments2Fl (write_output.f90:975-1041) of an oracle box, FP64."""
    pr, pb = o.pops(0), o.pops(1)
    a, b = pr.sum(-1), pb.sum(-1)
    j = (pr + pb) @ E
    ok = (o.flags() == 1) & (a + b > 0)
    vel = np.where(ok[..., None], j / (a + b)[..., None], 0.0)
    ph = np.where(ok, (a - b) / (a + b), 0.0)
    return vel, ph


DROPLET_VTK = DROPLET  # nIterVTK = 10 over 20 steps


def test_vtk_files_of_the_droplet_match_the_oracle(cpu_host, tmp_path):
    from lbcuda_amd.abi import default_config, INIT_DROPLET
    from oracle.pyoracle import OracleBox
    d = str(tmp_path / "vtk")
    write_case(d, DROPLET_VTK, nml="&lbcuda box=16, sigma=0.05, beta=0.7 /\n")
    r = run(cpu_host, d)
    assert r.returncode == 0, r.stdout + r.stderr
    names = sorted(os.listdir(os.path.join(d, "output")))
    want = [f"{v}_000000.{s:06d}.vti" for v in ("phase", "rho", "rhoB", "vel") for s in (0, 10, 20)]
    assert names == sorted(want)

    o = OracleBox(default_config(16, sigma=0.05, beta=0.7), 64)
    o.init_fields(INIT_DROPLET)
    for step in (0, 10, 20):
        if step:
            o.step(10)
        rr, rb = o.rho()
        vel, ph = oracle_moments(o)
        for var, name, ref, nc in (("rho", "rho1", rr, 1), ("rhoB", "rho2", rb, 1),
                                   ("phase", "phase", ph, 1), ("vel", "vel", vel, 3)):
            head, data, tail = read_vti(os.path.join(d, "output", f"{var}_000000.{step:06d}.vti"))
            assert 'WholeExtent="000001 000016 000001 000016 000001 000016"' in head
            assert f'Name="{name}"' in head and 'format="appended"' in head
            assert ('NumberOfComponents="3"' in head) == (nc == 3)
            assert data.size == nc * 16 ** 3
            assert "</AppendedData>" in tail and "</VTKFile" in tail
            if nc == 3:   # (3, nx, ny, nz) Fortran order = x-fastest nodes of 3 components
                got = data.reshape(3, 16, 16, 16, order="F").transpose(1, 2, 3, 0)
            else:
                got = data.reshape(16, 16, 16, order="F")
            np.testing.assert_allclose(got, ref.astype(np.float32), rtol=1e-6, atol=1e-9,
                                       err_msg=f"{var} step {step}")
    o.close()


def test_restart_files_continue_the_droplet(cpu_host, tmp_path):
    """restart_every writes the initType 4 files; a run started from them
    (initType 4) reproduces the uninterrupted run up to the float rounding of
    the files (the CPU implementation runs in FP64)."""
    d1 = str(tmp_path / "full")
    write_case(d1, DROPLET_VTK, nml="&lbcuda box=16, sigma=0.05, beta=0.7, restart_every=10 /\n")
    r = run(cpu_host, d1)
    assert r.returncode == 0, r.stdout + r.stderr
    rdir = os.path.join(d1, "restart_000010")
    files = sorted(os.listdir(rdir))
    assert files == sorted(["rhoR.init.bin", "rhoB.init.bin"] +
                           [f"pop{c}.{l:06d}.init.bin" for c in "RB" for l in range(19)])
    ext, rho_r = read_init_bin(os.path.join(rdir, "rhoR.init.bin"), (16, 16, 16))
    assert list(ext) == [1, 16, 1, 16, 1, 16]
    _, rho1, _ = read_vti(os.path.join(d1, "output", "rho_000000.000010.vti"))
    assert np.array_equal(rho_r, rho1.reshape(16, 16, 16, order="F"))

    d2 = str(tmp_path / "again")
    os.makedirs(d2)
    for f in files:
        shutil.copy(os.path.join(rdir, f), d2)
    write_case(d2, DROPLET_VTK.replace("20\t!niter", "10\t!niter").replace("1               ! init", "4 ! init"),
               nml="&lbcuda box=16, sigma=0.05, beta=0.7 /\n")
    r2 = run(cpu_host, d2)
    assert r2.returncode == 0, r2.stdout + r2.stderr
    m_full = [float(x) for x in r.stdout.split("final mass R,B:")[1].split()[:2]]
    m_again = [float(x) for x in r2.stdout.split("final mass R,B:")[1].split()[:2]]
    np.testing.assert_allclose(m_again, m_full, rtol=1e-6)
    for var in ("phase", "vel"):
        _, a, _ = read_vti(os.path.join(d1, "output", f"{var}_000000.000020.vti"))
        _, b, _ = read_vti(os.path.join(d2, "output", f"{var}_000000.000010.vti"))
        np.testing.assert_allclose(b, a, rtol=0, atol=2e-6, err_msg=var)


def rotation_zyx(yaw, pitch, roll):
    """Rotation matrix of yaw about z, then pitch about y, then roll about x
    (intrinsic), from the elementary rotations."""
    cz, sz, cy, sy, cx, sx = (math.cos(yaw), math.sin(yaw), math.cos(pitch), math.sin(pitch),
                              math.cos(roll), math.sin(roll))
    rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    return rz @ ry @ rx


def read_particle_vtk(path):
    lines = open(path).read().split("\n")
    assert lines[0] == "# vtk DataFile Version 2.0" and lines[2] == "ASCII" and lines[3] == "DATASET POLYDATA"
    n = int(lines[4].split()[1])
    out = {"POINTS": np.array([[float(v) for v in lines[5 + i].split()] for i in range(n)])}
    i = 5 + n
    while i < len(lines):
        if lines[i].startswith("VECTORS "):
            name = lines[i].split()[1]
            out[name] = np.array([[float(v) for v in lines[i + 1 + k].split()] for k in range(n)])
            i += n
        elif lines[i].startswith("LOOKUP_TABLE"):
            out["Type"] = [int(lines[i + 1 + k]) for k in range(n)]
            i += n
        i += 1
    return n, out


def test_particle_vtk_and_restart_state(cpu_host, tmp_path):
    """writeParticleVTK (write_output.f90:674-746): positions, velocities and,
    with rotation, the body axes q e_i q*.  At step 0 the axes are the columns
    of the z-y-x rotation built from input.xyz's angles (read_input_atom reads
    them into eul2q(e1, e3, e2), main.CUF:1284); the restart writes input.xyz
    whose angles rebuild the same axes, and particles.init.bin with the full
    state, from which the restarted run starts exactly."""
    d = str(tmp_path / "pv")
    deck = particle_deck(1.75).replace("100\t\t\t!vtk", "6\t\t\t!vtk")
    write_case(d, deck, XYZ, "&lbcuda box=32, sigma=0.02, beta=0.7, wetting='neutral', restart_every=6 /\n")
    r = run(cpu_host, d)
    assert r.returncode == 0, r.stdout + r.stderr
    n, v0 = read_particle_vtk(os.path.join(d, "output", "particle_000000.vtk"))
    assert n == 4 and v0["Type"] == [1, 2, 3, 4]
    rows = [ln.split()[1:] for ln in XYZ.strip().split("\n")[2:]]
    xyz = np.array([[float(t) for t in row] for row in rows])
    np.testing.assert_allclose(v0["POINTS"], xyz[:, 0:3], rtol=1e-5)
    np.testing.assert_allclose(v0["vel"], xyz[:, 3:6], atol=1e-7)
    for a in range(4):
        e1, e2, e3 = xyz[a, 6:9]
        rot = rotation_zyx(e1, e3, e2)
        for ax, name in enumerate(("Vectx", "Vecty", "Vectz")):
            np.testing.assert_allclose(v0[name][a], rot[:, ax], atol=2e-6, err_msg=f"{a} {name}")

    # restart at step 6: input.xyz angles rebuild the written axes
    _, v6 = read_particle_vtk(os.path.join(d, "output", "particle_000006.vtk"))
    rows = [ln.split()[1:] for ln in open(os.path.join(d, "restart_000006", "input.xyz")).read().strip().split("\n")[2:]]
    rx = np.array([[float(t) for t in row] for row in rows])
    np.testing.assert_allclose(rx[:, 0:3], v6["POINTS"], rtol=1e-5)
    for a in range(4):
        e1, e2, e3 = rx[a, 6:9]
        rot = rotation_zyx(e1, e3, e2)
        for ax, name in enumerate(("Vectx", "Vecty", "Vectz")):
            np.testing.assert_allclose(v6[name][a], rot[:, ax], atol=2e-6, err_msg=f"{a} {name}")

    d2 = str(tmp_path / "pv2")
    os.makedirs(d2)
    for f in os.listdir(os.path.join(d, "restart_000006")):
        shutil.copy(os.path.join(d, "restart_000006", f), d2)
    deck2 = deck.replace("12\t           \t!niter", "6\t           \t!niter").replace("6\t\t\t! init", "4\t\t\t! init")
    write_case(d2, deck2, nml="&lbcuda box=32, sigma=0.02, beta=0.7, wetting='neutral' /\n")
    r2 = run(cpu_host, d2)
    assert r2.returncode == 0, r2.stdout + r2.stderr
    assert "particles.init.bin] Using the saved state" in r2.stdout
    _, w0 = read_particle_vtk(os.path.join(d2, "output", "particle_000000.vtk"))
    for k in ("POINTS", "vel", "Vectx", "Vecty", "Vectz"):
        assert np.array_equal(w0[k], v6[k]), k


@pytest.mark.gpu
def test_gpu_host_restart_from_checkpoint_is_exact(tmp_path):
    """The host linked against liblbcuda_amd.so: restart_every also writes the
    library's checkpoint image (state.ckpt); an initType 4 run beside it loads
    the image and its output equals the uninterrupted run's bit for bit, spheres
    included."""
    if not os.path.exists("/opt/rocm/bin/amdflang"):
        pytest.skip("amdflang not available")
    print("fortran host on the GPU library: checkpoint restart", flush=True)
    exe = build_host(str(tmp_path), os.path.join(ROOT, "lbcuda_amd", "_lib"), "lbcuda_amd")
    d1 = str(tmp_path / "full")
    deck = particle_deck(1.75).replace("100\t\t\t!vtk", "6\t\t\t!vtk")
    nml = "&lbcuda box=32, sigma=0.02, beta=0.7, wetting='neutral', storage='f16', restart_every=6 /\n"
    write_case(d1, deck, XYZ, nml)
    r = run(exe, d1)
    assert r.returncode == 0, r.stdout + r.stderr
    rdir = os.path.join(d1, "restart_000006")
    assert os.path.exists(os.path.join(rdir, "state.ckpt"))
    d2 = str(tmp_path / "again")
    os.makedirs(d2)
    for f in os.listdir(rdir):
        shutil.copy(os.path.join(rdir, f), d2)
    deck2 = deck.replace("12\t           \t!niter", "6\t           \t!niter").replace("6\t\t\t! init", "4\t\t\t! init")
    write_case(d2, deck2, nml=nml.replace(", restart_every=6", ""))
    r2 = run(exe, d2)
    assert r2.returncode == 0, r2.stdout + r2.stderr
    assert "state.ckpt] Continuing from the saved state" in r2.stdout
    for var in ("rho", "rhoB", "phase", "vel"):
        a = open(os.path.join(d1, "output", f"{var}_000000.000012.vti"), "rb").read()
        b = open(os.path.join(d2, "output", f"{var}_000000.000006.vti"), "rb").read()
        assert a == b, var
    a = open(os.path.join(d1, "output", "particle_000012.vtk")).read()
    b = open(os.path.join(d2, "output", "particle_000006.vtk")).read()
    assert a == b


def test_diffdens_deck(cpu_host, tmp_path):
    """diffdens in lbcuda.nml selects the density-ratio model: the droplet
    start scales the blue density by densB (setup2, kernels_fluid_CG.CUF:163-169)
    and both colour masses are conserved."""
    d = str(tmp_path / "dd")
    write_case(d, DROPLET, nml="&lbcuda box=16, sigma=0.05, beta=0.7, diffdens=.true., densb=4.0 /\n")
    r = run(cpu_host, d)
    assert r.returncode == 0, r.stdout + r.stderr
    m0 = [float(x) for x in r.stdout.split("initial mass R,B:")[1].split()[:2]]
    m1 = [float(x) for x in r.stdout.split("final mass R,B:")[1].split()[:2]]
    d0 = str(tmp_path / "plain")
    write_case(d0, DROPLET, nml="&lbcuda box=16, sigma=0.05, beta=0.7 /\n")
    r0 = run(cpu_host, d0)
    p0 = [float(x) for x in r0.stdout.split("initial mass R,B:")[1].split()[:2]]
    assert m0[0] == p0[0] and abs(m0[1] - 4.0 * p0[1]) < 1e-9 * m0[1]
    assert np.allclose(m0, m1, rtol=1e-10)
