"""The Fortran driver (mpi-multi-process-version_amd/fortran/lbm_driver).

It replaces the reference's `program main`: namelist inputs in place of the
compile-time parameters of `commondata`, the same check lines
(natural-convection_2dim_mono.f90:801) and the same Tecplot file
(output_ASCII, natural-convection_2dim_mono.f90:818-843 and
natural-convection_3dim_mono.f90:1097-1121).

CPU tests cover the input checks and error paths (they stop before any GPU
call).  GPU tests run small cavities and compare the printed residuals, the
Tecplot file and the FP64 binary dump with the oracle on the same inputs.
"""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "mpi-multi-process-version_amd", "fortran", "lbm_driver")
NUM = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[EeDd][-+]?\d+)?")


def _nml(path, **kw):
    def fmt(v):
        if isinstance(v, str):
            return f"'{v}'"
        return repr(v)
    body = ", ".join(f"{k}={fmt(v)}" for k, v in kw.items())
    with open(path, "w") as fh:
        fh.write(f"&lbm\n {body}\n/\n")
    return path


def _run(tmp_path, timeout=120, **kw):
    if not os.path.exists(DRIVER):
        pytest.fail("Fortran driver not built: run __graft_entry__.build() first")
    nml = _nml(str(tmp_path / "in.nml"), **kw)
    return subprocess.run([DRIVER, nml], cwd=str(tmp_path), capture_output=True,
                          text=True, timeout=timeout)


# --------------------------------------------------------------- CPU tests

def test_fortran_module_binds_every_c_entry_point():
    """fortran/lbm_capi_mod.f90 has a bind(C) interface for every function
    declared in include/*.h, under the same name."""
    decl = re.compile(r"^\s*(?:int32_t|lbm_handle|const char \*)\s*\*?\s*(lbm_\w+)\s*\(", re.M)
    names = set()
    for h in ("lbm_capi.h", "lbm_nn.h"):
        names |= set(decl.findall(open(os.path.join(ROOT, "include", h)).read()))
    mod = open(os.path.join(ROOT, "mpi-multi-process-version_amd", "fortran",
                            "lbm_capi_mod.f90")).read()
    bound = set(re.findall(r"bind\(C,\s*name='(lbm_\w+)'\)", mod))
    assert len(names) > 40
    assert names == bound, (sorted(names - bound), sorted(bound - names))



def test_rejects_bad_grid_before_touching_the_gpu(tmp_path):
    r = _run(tmp_path, dim=2, nx=3, Ra=1e4)
    assert r.returncode == 1
    assert "lbm_create failed: status 1 (invalid argument)" in r.stdout


def test_rejects_unstable_relaxation_time(tmp_path):
    # tau_f = 3 nu + 1/2 with nu = Ma H sqrt(Pr)/sqrt(3 Ra): Ma tiny -> tau_f ~ 1/2
    r = _run(tmp_path, dim=2, nx=33, Ra=1e4, Ma=1e-9)
    assert r.returncode == 1 and "invalid argument" in r.stdout


def test_rejects_bad_inputs(tmp_path):
    r = _run(tmp_path, dim=4)
    assert r.returncode == 2 and "dim must be 2 or 3" in r.stdout
    r = _run(tmp_path, dim=2, output="vtk")
    assert r.returncode == 2 and "output must be" in r.stdout
    r = _run(tmp_path, dim=2, check_every=0)
    assert r.returncode == 2 and "check_every" in r.stdout
    r = subprocess.run([DRIVER, str(tmp_path / "missing.nml")], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2 and "not found" in r.stdout
    bad = tmp_path / "garbage.nml"
    bad.write_text("&lbm\n nx = 'x'\n/\n")
    r = subprocess.run([DRIVER, str(bad)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "cannot read namelist" in r.stdout


# --------------------------------------------------------------- GPU tests

def _check_lines(stdout):
    rows = []
    for line in stdout.splitlines():
        tok = NUM.findall(line)
        if len(tok) == 3 and line.split()[0].isdigit() and len(line.split()) == 3:
            rows.append((int(tok[0]), float(tok[1]), float(tok[2])))
    return rows


def _read_tecplot(path, dim):
    with open(path) as fh:
        lines = fh.read().splitlines()
    header = lines[:3]
    data = np.array([[float(t) for t in NUM.findall(ln)] for ln in lines[3:]])
    return header, data, lines[3:]


def _read_binary(path):
    with open(path, "rb") as fh:
        dim, nx, ny, nz = np.frombuffer(fh.read(16), np.int32)
        itc = int(np.frombuffer(fh.read(8), np.int64)[0])
        nu = np.frombuffer(fh.read(16), np.float64)
        rest = np.frombuffer(fh.read(), np.float64)
    nn = nx * ny * nz
    axes = [nx, ny] + ([nz] if dim == 3 else [])
    pos = sum(axes)
    coords, o = [], 0
    for a in axes:
        coords.append(rest[o:o + a])
        o += a
    names = ["rho", "u", "v"] + (["w"] if dim == 3 else []) + ["T"]
    fields = {k: rest[pos + i * nn: pos + (i + 1) * nn] for i, k in enumerate(names)}
    assert rest.size == pos + len(names) * nn
    return dict(dim=dim, n=(nx, ny, nz), itc=itc, nu=nu, coords=coords, fields=fields)


def _oracle_run(oracle_mod, dim, n, Ra, steps, check_every):
    o = oracle_mod.Oracle(dim, n, Ra)
    o.init_rest()
    checks = []
    done = 0
    while done < steps:
        o.step(check_every)
        done += check_every
        checks.append((done,) + tuple(o.check()))
    return o, checks


@pytest.mark.gpu
def test_driver_2d_fp64_matches_oracle(tmp_path, oracle_mod):
    print("\n[driver] 2D 33^2 Ra=1e4 FP64 4000 steps", flush=True)
    r = _run(tmp_path, dim=2, nx=33, Ra=1e4, precision=1, itc_max=4000, output="both")
    print(r.stdout[-1500:], r.stderr[-500:], flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Successfully: DNS completed!" in r.stdout
    o, checks = _oracle_run(oracle_mod, 2, 33, 1e4, 4000, 2000)
    got = _check_lines(r.stdout)
    assert [c[0] for c in got] == [2000, 4000]
    assert got[0][1] == 1.0 and got[0][2] == 1.0          # first check is always 1
    for g, c in zip(got, checks):
        assert abs(g[1] - c[1]) <= 1e-12 * c[1] and abs(g[2] - c[2]) <= 1e-12 * c[2]
    # FP64 binary dump: the oracle's fields, operation for operation
    b = _read_binary(str(tmp_path / "lbm_fields-4000.bin"))
    assert b["dim"] == 2 and b["n"] == (33, 33, 1) and b["itc"] == 4000
    rho, u, v, w, T = o.fields()
    for k, ref in (("rho", rho), ("u", u), ("v", v), ("T", T)):
        err = np.abs(b["fields"][k] - ref.ravel()).max() / max(np.abs(ref).max(), 1e-300)
        assert err <= 1e-13, (k, err)
    x = y = o.coords()   # the same node coordinates on every axis
    assert np.array_equal(b["coords"][0], x) and np.array_equal(b["coords"][1], y)
    nh, nc = o.nusselt()
    assert abs(b["nu"][0] - nh) <= 1e-12 and abs(b["nu"][1] - nc) <= 1e-12
    # Tecplot file: header and POINT order of output_ASCII, values to the
    # printed precision (e11.4 coordinates, e13.6 fields)
    header, data, raw = _read_tecplot(str(tmp_path / "MRTcavity-4000.dat"), 2)
    assert header[0].strip() == 'TITLE="thermal convective flows"'
    assert header[1].strip() == 'VARIABLES="X" "Y" "U" "V" "T" "rho"'
    assert header[2] == "ZONE I=    33  J=    33 F=POINT"
    assert data.shape == (33 * 33, 6)
    assert re.fullmatch(r" ( \d\.\d{4}E[-+]\d\d ){2}(( |-)\d\.\d{6}E[-+]\d\d ){4}", raw[0]), raw[0]
    # rounding of Ew.d is at most half a unit of the d-th digit of a mantissa
    # in [0.1, 1): 5e-4 (e11.4) and 5e-6 (e13.6) relative to the value
    X, Y = np.meshgrid(x, y)
    for col, ref, rel in ((0, X, 5e-4), (1, Y, 5e-4), (2, u, 5e-6), (3, v, 5e-6),
                          (4, T, 5e-6), (5, rho, 5e-6)):
        tol = rel * np.abs(ref).ravel() * (1 + 1e-9) + 1e-13 * np.abs(ref).max()
        assert np.all(np.abs(data[:, col] - ref.ravel()) <= tol), col
    o.close()


@pytest.mark.gpu
def test_driver_3d_fp32_matches_oracle(tmp_path, oracle_mod):
    print("\n[driver] 3D 17^3 Ra=1e5 FP32 2000 steps", flush=True)
    r = _run(tmp_path, dim=3, nx=17, Ra=1e5, precision=0, itc_max=2000,
             check_every=1000, output="both")
    print(r.stdout[-1500:], r.stderr[-500:], flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    o, checks = _oracle_run(oracle_mod, 3, 17, 1e5, 2000, 1000)
    got = _check_lines(r.stdout)
    assert [c[0] for c in got] == [1000, 2000] and got[0][1:] == (1.0, 1.0)
    assert abs(got[1][1] - checks[1][1]) <= 1e-4 * checks[1][1]
    assert abs(got[1][2] - checks[1][2]) <= 1e-4 * checks[1][2]
    b = _read_binary(str(tmp_path / "lbm_fields-2000.bin"))
    assert b["dim"] == 3 and b["n"] == (17, 17, 17) and b["itc"] == 2000
    rho, u, v, w, T = o.fields()
    vscale = max(np.abs(u).max(), np.abs(v).max(), np.abs(w).max())
    for k, ref, scale in (("rho", rho, 1.0), ("u", u, vscale), ("v", v, vscale),
                          ("w", w, vscale), ("T", T, np.abs(T).max())):
        err = np.abs(b["fields"][k] - ref.ravel()).max() / scale
        assert err <= 1e-5, (k, err)
    header, data, raw = _read_tecplot(str(tmp_path / "side-heated convection17.dat"), 3)
    assert header[0].strip() == 'TITLE="side-heated convection "'
    assert header[1].strip() == 'VARIABLES="X" "Y" "Z" "U" "V" "W" "RHO" "T"'
    assert header[2] == "ZONE I=    17  J=    17 K=    17 F=POINT"
    # format 100 applies e11.4 to X and Y only; Z and the fields are e13.6
    assert re.fullmatch(r" ( \d\.\d{4}E[-+]\d\d ){2}(( |-)\d\.\d{6}E[-+]\d\d ){6}", raw[0]), raw[0]
    assert data.shape == (17 ** 3, 8)
    o.close()


@pytest.mark.gpu
def test_driver_switches_from_fp32_to_fp64(tmp_path, oracle_mod):
    """switch_eps: the FP32 run hands its state to FP64 once both residuals
    are at or below 1e-4, then stops on the reference's 1e-8 rule; the
    result is the FP64 steady state (the oracle's, run to the same rule)."""
    print("\n[driver] 2D 33^2 Ra=1e4 FP32 -> FP64 at 1e-4, to 1e-8", flush=True)
    r = _run(tmp_path, dim=2, nx=33, Ra=1e4, precision=0, switch_eps=1e-4, epsU=1e-8,
             epsT=1e-8, itc_max=400000, output="binary")
    print(r.stdout[-800:], r.stderr[-300:], flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "switched to FP64 at itc =" in r.stdout
    got = _check_lines(r.stdout)
    assert got[-1][1] <= 1e-8 and got[-1][2] <= 1e-8
    at = int(re.search(r"switched to FP64 at itc =\s*(\d+)", r.stdout).group(1))
    row = [g for g in got if g[0] == at][0]
    assert row[1] <= 1e-4 and row[2] <= 1e-4 and at < got[-1][0]
    o = oracle_mod.Oracle(2, 33, 1e4)
    o.init_rest()
    itc, eu, et = o.run(1e-8, 1e-8, 400000)
    b = _read_binary(str(tmp_path / f"lbm_fields-{got[-1][0]}.bin"))
    print(f"  switched at {at}, stopped at {got[-1][0]} (FP64 oracle: {itc})", flush=True)
    nh, _ = o.nusselt()
    assert abs(b["nu"][0] / nh - 1) <= 1e-6
    rho, u, v, w, T = o.fields()
    vs = np.abs(np.stack([u, v])).max()
    assert np.abs(b["fields"]["u"] - u.ravel()).max() / vs <= 1e-5
    assert np.abs(b["fields"]["T"] - T.ravel()).max() <= 1e-5
    o.close()
