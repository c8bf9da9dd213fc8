"""Command-line tool: flags, defaults and messages of src/cli.rs:13-329 that
need no GPU (help, version, argument errors).  The runs themselves are in
test_sim_gpu.py."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "johansen-null-eigenspectra_amd", "bin", "johansen-null-eigenspectra")


def run(*args):
    if not os.path.exists(CLI):
        pytest.fail(f"{CLI} not built (run __graft_entry__.build())")
    p = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=60)
    return p.returncode, p.stdout, p.stderr


def test_help_lists_reference_flags_and_models():
    rc, out, _ = run("--help")
    assert rc == 0 and out.startswith(f"Usage: {CLI} [OPTIONS]")
    for line in ["  --steps <int>        number of simulation steps (default: 10,000)",
                 "  --runs <int>         number of runs per model (default: 10,000,000)",
                 "  --dim-start <int>    starting matrix dimension (default: 1)",
                 "  --dim-end <int>      ending matrix dimension (default: 12)",
                 "  --dim <int>          run a single dimension (sets start and end to the same value)",
                 "  --model <list>       comma separated list of model numbers to compute (default: 0,1,2,3,4)",
                 "  --quiet              suppress progress output",
                 "  4: Intercept, trend, intercept and trend not fully explained by cointegration",
                 f"  {CLI} --model 0,2 --runs 100,000"]:
        assert line in out.splitlines(), line
    assert run("-h")[1] == out


def test_version():
    rc, out, _ = run("--version")
    assert rc == 0 and out.startswith("johansen-null-eigenspectra version ")


@pytest.mark.parametrize("args,err", [
    (["--bogus"], "Error: unknown argument '--bogus'\nUse --help to see available options\n"),
    (["--steps"], "Error: --steps parameter requires a value\n"),
    (["--runs", "1e6"], "Error: --runs parameter must be a positive integer (commas allowed, e.g., 1,000,000)\n"),
    (["--dim-start", "5", "--dim-end", "3"], "Error: dimension start (5) cannot be greater than end (3)\n"),
    (["--steps", "0"], "Error: steps, runs, and dimension parameters must be greater than 0\n"),
    (["--model", "0,7"], "Error: 無效的模型代號: 7\n"),
    (["--model", ","], "Error: 模型列表不可為空\n"),
    (["--threads", "100000"], None),
    # the engine's limits, checked before any device work
    (["--dim", "33"], "Error: dimension 33 is above the supported maximum of 32\n"),
    (["--steps", "1,000,000,000,000"], "Error: steps must be between 2 and 1,073,741,824\n"),
    (["--steps", "1"], "Error: steps must be between 2 and 1,073,741,824\n"),
    (["--runs", "4,294,967,296"], "Error: runs must be at most 4,294,967,295 (seeds are u32)\n"),
])
def test_argument_errors(args, err):
    rc, out, e = run(*args)
    assert rc == 0 and out == ""
    if err is None:  # more than twice the logical cores is refused without a prompt
        assert "Error: excessive thread count rejected to avoid system overload." in e
    else:
        assert e == err
