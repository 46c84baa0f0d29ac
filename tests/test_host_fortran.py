"""The Fortran host program (host/lbcuda_main.f90) keeps the reference's
input.dat / input.xyz semantics (readParameters main.CUF:409-496,
read_input_atom :1230-1350) and drives the C ABI.  On the CPU it is linked
against the oracle-backed implementation of the same ABI
(oracle/liblbm_cpu_abi.so); on a GPU, against liblbcuda_amd.so."""
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
