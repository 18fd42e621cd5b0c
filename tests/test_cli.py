"""The lucasntt command line: the reference's INI keys, arguments, files, resume and
verdicts (main_process.cu:64-196, lucas_prp.cu:137-346)."""
import os
import signal
import subprocess
import time

import pytest

from helpers import ROOT, OracleTest, golden

CLI = os.path.join(ROOT, "cuda_amd", "cli", "lucasntt")


def run(args, cwd, env=None, stdin=None, timeout=600):
    e = dict(os.environ)
    if env:
        e.update(env)
    # stdin=None: a closed stdin (end of input at once)
    kw = {"input": stdin} if stdin is not None else {"stdin": subprocess.DEVNULL}
    return subprocess.run([CLI] + [str(a) for a in args], cwd=cwd, env=e, capture_output=True,
                          text=True, timeout=timeout, **kw)


def dry(args, cwd, stdin=""):
    r = run(args, cwd, env={"LNT_DRY_RUN": "1"}, stdin=stdin)
    line = [l[l.index("DRYRUN"):] for l in r.stdout.splitlines() if "DRYRUN" in l]
    return r, (dict(kv.split("=", 1) for kv in line[0].split()[1:]) if line else None)


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(CLI):
        subprocess.run(["make", "-C", os.path.dirname(CLI)], check=True, capture_output=True)


def test_defaults_without_ini(tmp_path):
    r, c = dry([], tmp_path)
    assert r.returncode == 0 and "Welcome to LucasNTT" in r.stdout
    assert c["exponent"] == "9689" and c["iterations"] == "9687" and c["PRP"] == "0"
    assert c["dump"] == "LucasNTT_results.dat" and c["stats"] == "LucasNTT_results.txt"
    assert c["fromIteration"] == "1" and c["loops"] == "500"


def test_ini_keys_and_argument_precedence(tmp_path):
    (tmp_path / "LucasNTT.ini").write_text(
        "## comment\n# Exponent = 1\nExponent = 86243\nPRP = 1\nIterations = 1000\n"
        "PerformanceLoops = 20\nStats = mystats.txt\nDump = myres.dat\ndumpResidue = 1\n"
        "Pool = FFT_2SMEM\nKernels = fftr__r2_8to4k;m2\nfft = 64:64\n")
    r, c = dry([], tmp_path, stdin="Y\n")
    assert c["exponent"] == "86243" and c["PRP"] == "1" and c["iterations"] == "1000"
    assert c["stats"] == "mystats.txt" and c["dump"] == "myres.dat" and c["dumpResidue"] == "1"
    assert c["pool"] == "FFT_2SMEM" and c["kernels"] == ";fftr__r2_8to4k;m2;" and c["fft"] == "64:64"
    assert c["loops"] == "20"
    # positional arguments win: exponent, iterations, output base, fft
    r, c = dry([9689, 500, "run1", "512"], tmp_path)
    assert c["exponent"] == "9689" and c["iterations"] == "500"
    assert c["dump"] == "run1.dat" and c["stats"] == "run1.txt" and c["env"] == "run1.env"
    # iterations beyond the full test are clamped (main_process.cu:116-117); PRP full = p
    r, c = dry([9689, 10 ** 6], tmp_path, stdin="Y\n")  # the INI fft (64:64) is not 9689's length
    assert c["iterations"] == "9689"


def test_non_secured_length_asks_for_confirmation(tmp_path):
    r, c = dry([86243, 100, "x", "2K"], tmp_path, stdin="N\n")
    assert r.returncode == 1 and "too small" in r.stdout and "Process cancelled by user" in r.stdout
    r, c = dry([86243, 100, "x", "8K"], tmp_path, stdin="")
    assert r.returncode == 1 and "not optimal" in r.stdout
    r, c = dry([86243, 100, "x", "8K"], tmp_path, stdin="maybe\ny\n")
    assert r.returncode == 0 and c["fft"] == "8K"
    r, c = dry([86243, 100, "x", "4K"], tmp_path)
    assert r.returncode == 0


def test_confirm_length_key_answers_without_stdin(tmp_path):
    """ConfirmLength answers the length question of main_process.cu:137-141 for scripted
    runs: stdin is never read (it is closed here), Y proceeds, N cancels."""
    (tmp_path / "LucasNTT.ini").write_text("ConfirmLength = Y\n")
    r, c = dry([86243, 100, "x", "8K"], tmp_path, stdin=None)
    assert r.returncode == 0 and c["fft"] == "8K" and c["confirmLength"] == "Y"
    assert "not optimal" in r.stdout and "(ConfirmLength)" in r.stdout
    (tmp_path / "LucasNTT.ini").write_text("ConfirmLength = no\n")
    r, c = dry([86243, 100, "x", "2K"], tmp_path, stdin=None)
    assert r.returncode == 1 and "Process cancelled by user" in r.stdout
    (tmp_path / "LucasNTT.ini").write_text("ConfirmLength = maybe\n")
    r, c = dry([86243, 100, "x", "2K"], tmp_path, stdin=None)
    assert r.returncode == 1 and "ConfirmLength must be Y or N" in r.stdout
    # without the key, a closed stdin still means "N"
    (tmp_path / "LucasNTT.ini").write_text("DumpIterations = 250\n")
    r, c = dry([86243, 100, "x", "8K"], tmp_path, stdin=None)
    assert r.returncode == 1 and "Process cancelled by user" in r.stdout
    r, c = dry([], tmp_path)
    assert c["dumpIterations"] == "250" and c["confirmLength"] == "ask"


def test_resume_reads_the_stats_file(tmp_path):
    (tmp_path / "LucasNTT_results.txt").write_text(
        "Exponent = 86243\nConclusion = Partial\nPRP = 0\nIterations = 86241\nfft = 4096\n"
        "n = 4096\nlog2_n = 12\nlog3_n = 0\nlog5_n = 0\nPool = LNT_1WG\nInterrupted = 1\n"
        "currentIteration = 12345\nPer iteration (ms) = 0.1\nTotal time (sec) = 1\n")
    r, c = dry([], tmp_path)
    assert "Last execution was interrupted" in r.stdout
    assert c["wasInterrupted"] == "1" and c["exponent"] == "86243" and c["fromIteration"] == "12345"
    assert c["fft"] == "4096"


def test_invalid_length_string(tmp_path):
    r, c = dry([86243, 100, "x", "7K"], tmp_path)
    assert r.returncode == 1 and "Invalid NTT length" in r.stdout


# ---------------------------------------------------------------- on the GPU

@pytest.mark.gpu
def test_full_ll_of_a_mersenne_prime(tmp_path):
    (tmp_path / "LucasNTT.ini").write_text("PerformanceLoops = 10\ndumpEnvironment = 1\n")
    r = run([9689], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "M(9689) is prime!" in r.stdout
    stats = (tmp_path / "LucasNTT_results.txt").read_text()
    keys = [l.split(" = ")[0] for l in stats.splitlines() if " = " in l]
    assert keys[:9] == ["Exponent", "Conclusion", "PRP", "Iterations", "fft", "n", "log2_n",
                        "log3_n", "log5_n"]
    assert "Conclusion = Prime" in stats and "Interrupted = 0" in stats
    assert "Res64 = 0000000000000000" in stats
    assert not (tmp_path / "LucasNTT_results.dat").exists()  # full tests do not dump
    assert (tmp_path / "LucasNTT_results.env").exists()


@pytest.mark.gpu
def test_prp_and_composite_verdicts(tmp_path):
    (tmp_path / "LucasNTT.ini").write_text("PerformanceLoops = 10\nPRP = 1\n")
    r = run([9689], tmp_path)
    assert "M(9689) is probably prime!" in r.stdout
    assert "Conclusion = Probably Prime" in (tmp_path / "LucasNTT_results.txt").read_text()
    r = run([9697, 0, "c"], tmp_path)
    assert "M(9697) is composite!" in r.stdout
    assert "Conclusion = Composite" in (tmp_path / "c.txt").read_text()


@pytest.mark.gpu
def test_partial_run_dumps_the_reference_residue_format(tmp_path):
    (tmp_path / "LucasNTT.ini").write_text("PerformanceLoops = 10\n")
    r = run([86243, 1000, "part"], tmp_path)
    assert r.returncode == 0 and "1000 loops performed with M(86243)" in r.stdout
    raw = (tmp_path / "part.dat").read_bytes()
    assert len(raw) == 4096 * 19
    lines = raw.decode().splitlines()
    assert all(len(l) == 18 and l.startswith("0x") and l[2:] == l[2:].upper() for l in lines)
    o = OracleTest(86243, 4096)
    o.iterate(1000)
    assert [int(l, 16) for l in lines] == o.digits()
    stats = (tmp_path / "part.txt").read_text()
    assert "Conclusion = Partial" in stats and f"Res64 = {golden(86243, 'LL')[1000]:016X}" in stats


@pytest.mark.gpu
def test_safety_dump_every_dump_iterations(tmp_path):
    """The periodic safety dump (lucas_prp.cu:207-214) runs before the iteration whose
    count reaches the interval: with DumpIterations = 1000 a full LL of M9689 (9,687
    iterations) dumps 9 times, the last one the residue after 8,999 iterations, with the
    stats file of that moment.  A full run writes no final .dat (and dumpStats = 0 no
    final stats), so the files left are the last safety dump's."""
    from cuda_amd import secured_log2_length
    (tmp_path / "LucasNTT.ini").write_text("PerformanceLoops = 0\nDumpIterations = 1000\n"
                                           "dumpStats = 0\n")
    r = run([9689, 0, "d"], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "M(9689) is prime!" in r.stdout
    assert r.stdout.count("Now dumping intermediate residue for safety reasons") == 9
    lines = (tmp_path / "d.dat").read_text().splitlines()
    o = OracleTest(9689, 1 << secured_log2_length(9689))
    o.iterate(8999)
    assert [int(l, 16) for l in lines] == o.digits()
    stats = (tmp_path / "d.txt").read_text()
    assert "currentIteration = 9000" in stats and "Interrupted = 1" in stats


@pytest.mark.gpu
def test_interrupt_and_resume_equals_uninterrupted(tmp_path):
    """SIGINT mid-run dumps residue and stats (Interrupted = 1); the next start reloads
    them and finishes with the verdict of an uninterrupted run."""
    (tmp_path / "LucasNTT.ini").write_text("PerformanceLoops = 10\n")
    p = subprocess.Popen([CLI, "86243"], cwd=tmp_path, stdout=subprocess.PIPE,
                         stderr=subprocess.STDOUT, text=True)
    out = ""
    while "From iteration" not in out:
        out += p.stdout.readline()
    time.sleep(1.0)
    p.send_signal(signal.SIGINT)
    out += p.communicate(timeout=300)[0]
    stats = (tmp_path / "LucasNTT_results.txt").read_text()
    if "Interrupted = 1" not in stats:
        pytest.skip("the run finished before the interrupt arrived")
    assert (tmp_path / "LucasNTT_results.dat").exists()
    r = run([], tmp_path)
    assert "Previous residue successfully loaded!" in r.stdout, r.stdout
    assert "M(86243) is prime!" in r.stdout
    assert not (tmp_path / "LucasNTT_results.dat").exists()  # LoadResidue removes the dump


@pytest.mark.gpu
def test_tuning_persists_in_the_tune_file(tmp_path):
    """A run that tunes writes LucasNTT.tune; a later run without tuning reads it and ends
    on the same residue."""
    (tmp_path / "LucasNTT.ini").write_text("PerformanceLoops = 10\n")
    r = run([200003, 300], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = (tmp_path / "LucasNTT.tune").read_text().splitlines()
    assert any(l.split()[:2] == ["0", "8192"] for l in lines), lines
    first = (tmp_path / "LucasNTT_results.txt").read_text()
    (tmp_path / "LucasNTT_results.txt").unlink()
    (tmp_path / "LucasNTT.ini").write_text("PerformanceLoops = 0\n")
    r = run([200003, 300], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    res = [l for l in first.splitlines() if l.startswith("Res64")]
    assert res and res[0] in (tmp_path / "LucasNTT_results.txt").read_text()
