"""The command-line driver nbody_amd/bin/nbody (flags and output of
src/nbody.cu:735-1122). CPU tests check argument handling with no GPU in
sight; the GPU test runs the reference CI's smoke command
(.github/script.sh: --bodies 4 --cycle-after 1 --iterations 1 --verbose) and
parses every line in the format of src/nbody.cu:1033-1050."""
import os
import re
import subprocess

import pytest

from conftest import ROOT

EXE = os.path.join(ROOT, "nbody_amd", "bin", "nbody")
NAMES = ["HIP_AOS", "HIP_LDS", "HIP_DPP", "HIP_SGPR", "HIP_Symmetric", "HIP", "HIP_Host"]

# one result line: "<name>: <ms> ms = <rate>x10^9 interactions/s (<gflops> GFLOPS) (Rel. error: <E>)"
LINE = re.compile(r"^\s*(?P<name>\w+):\s+(?P<ms>[\d.]+) ms =\s+(?P<rate>[\d.]+)x10\^(?P<exp>[69]) interactions/s "
                  r"\(\s*(?P<gflops>[\d.]+) GFLOPS\)(?: \(Rel\. error: (?P<err>[-+\d.E]+)\))?$")


def run(*args, timeout=120):
    return subprocess.run([EXE, *args], capture_output=True, text=True, timeout=timeout)


def test_cli_exists():
    assert os.access(EXE, os.X_OK), "run python -m nbody_amd.build"


@pytest.mark.parametrize("args,msg", [
    (["--bodies", "0"], "Requested number of bodies less than 1"),
    (["--bodies", "x"], "Couldn't parse integer argument for '--bodies'"),
    (["--seed", "0"], "Seed must be nonzero integer value"),
    (["--iterations", "0"], "Requested number of iterations less than 1"),
    (["--cycle-after", "0"], "Requested cycle size less than 1"),
    (["--algorithm", "GPU_Shared"], "Invalid algorithm name 'GPU_Shared'"),
    (["--algorithm", "99"], "Invalid algorithm index '99'"),
    (["--ic", "cube"], "--ic must be reference, ball or plummer"),
    (["--golden", "cpu"], "--golden must be fp64 or cpu_soa"),
])
def test_cli_rejects_bad_arguments(args, msg):
    r = run(*args)
    assert r.returncode == 1 and msg in r.stderr, (r.returncode, r.stderr)


def test_cli_help_and_list():
    r = run("--help")
    assert r.returncode == 1 and "--bodies=<N>" in r.stderr and "--algorithm" in r.stderr and "--golden" in r.stderr
    r = run("--list")
    assert r.returncode == 0
    listed = re.findall(r"^\s+(\d+) - (\w+)", r.stdout, re.M)
    assert [n for _, n in listed] == NAMES
    assert [int(i) for i, _ in listed] == list(range(len(NAMES)))


def test_line_format_matches_reference_printf():
    # the reference's own printf format (src/nbody.cu:1033-1050), rendered here
    line = "%*s: %8.2f ms = %8.3fx10^9 interactions/s (%9.2f GFLOPS)" % (13, "HIP", 1.94, 4981.031, 99620.62)
    line += " (Rel. error: %E)" % 3.2e-6
    m = LINE.match(line)
    assert m and m.group("name") == "HIP" and m.group("err") == "3.200000E-06"


@pytest.mark.gpu
def test_cli_reference_ci_smoke():
    r = run("--bodies", "4", "--cycle-after", "1", "--iterations", "1", "--verbose", timeout=300)
    assert r.returncode == 0, r.stderr
    assert "Running simulation with 4096 particles, crosscheck enabled" in r.stderr
    lines = [l for l in r.stdout.replace("\r", "\n").splitlines() if l.strip()]
    got = {}
    for l in lines:
        m = LINE.match(l)
        assert m, l
        got[m.group("name")] = m
    assert sorted(got) == sorted(NAMES)
    for name, m in got.items():
        assert float(m.group("ms")) > 0
        # componentwise relError of the reference (signed; tiny golden components inflate it)
        assert abs(float(m.group("err"))) < 5e-2, (name, m.group("err"))


@pytest.mark.gpu
def test_cli_ball_and_plummer_run():
    for ic in ("ball", "plummer"):
        r = run("--bodies", "2", "--iterations", "3", "--algorithm", "HIP", "--ic", ic, "--verbose", timeout=300)
        assert r.returncode == 0, r.stderr
        assert len([l for l in r.stdout.replace("\r", "\n").splitlines() if LINE.match(l)]) == 3


@pytest.mark.gpu
def test_cli_cpu_soa_golden():
    # --golden=cpu_soa: the reference's own crosscheck golden (src/nbody.cu:311-321,
    # CPU_SOA), so "Rel. error" is the figure the reference's logs print.
    errs = {}
    for golden in ("fp64", "cpu_soa"):
        r = run("--bodies", "2", "--iterations", "2", "--algorithm", "HIP", "--golden", golden, "--verbose",
                timeout=300)
        assert r.returncode == 0, r.stderr
        if golden == "cpu_soa":
            assert "crosscheck enabled (CPU_SOA golden)" in r.stderr
        ms = [LINE.match(l) for l in r.stdout.replace("\r", "\n").splitlines() if LINE.match(l)]
        assert len(ms) == 2 and all(m.group("err") is not None for m in ms)
        errs[golden] = [abs(float(m.group("err"))) for m in ms]
    # both goldens agree with the GPU force to the FP32 rounding of the sum
    # (componentwise, so small components inflate it; the FP64 figure bounds
    # how far the FP32 golden itself can sit from the GPU's result)
    for e64, esoa in zip(errs["fp64"], errs["cpu_soa"]):
        assert esoa < 5e-2 and e64 < 5e-2, errs
