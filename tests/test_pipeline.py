"""Timing rows and summaries in the artifact's pipeline formats (run.sh,
aggregate.sh, model/compute_gamma.py), checked against the reference's own
recorded results: the aggregator must rebuild its MI250X HIP summary.csv byte
for byte from its run logs, and the machine constant its gamma_single_gpu.csv
files from their summaries (tests/golden/make_pipeline_fixture.py wrote the
fixtures)."""
import io
import os
import stat
import sys

import pytest

import h3d_pipeline as P

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return f.read()


def test_aggregate_rebuilds_reference_summary():
    recs = P.parse_run_log(golden("hip_mi250x_runlog.txt").splitlines())
    assert len(recs) == 315  # 35 jobs x 9 variants, 5 repetitions of 63 cases
    assert "\n".join(P.summarize(recs)) + "\n" == golden("hip_mi250x_summary.csv")


def test_headers_match_reference():
    assert P.TIMINGS_HEADER + "\n" == golden("hip_mi250x_timings_header.txt")
    assert P.SUMMARY_HEADER == golden("hip_mi250x_summary.csv").splitlines()[0]


def test_sim_time_matches_recorded_dt():
    # 'N=... DT=...' header lines of the reference's results_*/run.log
    recorded = {512: "0.0001679668", 640: "0.0001074146", 768: "7.45546e-05",
                896: "5.475441e-05", 1024: "4.190964e-05", 1152: "3.310659e-05",
                1280: "2.681168e-05"}
    assert {n: P.sim_time(n) for n in P.N_LIST} == recorded


def test_parse_rules():
    log = """\
>>> Running: ./1-hip  gpus=1  mask=0  N=64
Estimated timesteps:     ~500
Time = 1.5 seconds
>>> Running: 2-omp  gpus=2  mask=0,2  N=64
501 timesteps complete
Time = 9.0 seconds
>>> Running: 1-hip-h3d  gpus=1  mask=0  N=64
Estimated timesteps:     ~500
Time=3.0 seconds
501 timesteps complete
Time = 2.5 seconds
>>> Running: 1-hip  gpus=1  mask=0  N=64
Estimated timesteps:     ~500
>>> Running: 1-hip  gpus=1  mask=0  N=64
Time = 0.5 seconds
"""
    recs = P.parse_run_log(log.splitlines())
    # './' dropped; a variant that is not ^[0-9]+-hip is skipped; a Time
    # line without a spaced '=' does not close the run; a run with no Time
    # line is dropped; steps: 'complete' over the estimate, '' without either
    assert recs == [("1-hip", "1", "64", "1.5", "500"),
                    ("1-hip-h3d", "1", "64", "2.5", "501"),
                    ("1-hip", "1", "64", "0.5", "")]
    rows = P.summarize(recs)
    # steps keeps the last non-empty value; median of an even count is the
    # mean of the middle pair
    assert rows[1:] == ["1-hip,1,64,2,0,0.500000,1.000000,1.500000,100.0,500",
                        "1-hip-h3d,1,64,1,0,2.500000,2.500000,2.500000,0.0,501"]


def test_summary_order_and_median():
    recs = [("2-hip", "2", "640", "4", "500"), ("2-hip", "2", "96", "1", "500"),
            ("10-hip", "1", "512", "3", "501"), ("2-hip", "2", "640", "1", "500"),
            ("2-hip", "2", "640", "2", "500")]
    rows = P.summarize(recs)
    assert [r.split(",")[:3] for r in rows[1:]] == [["10-hip", "1", "512"],
                                                    ["2-hip", "2", "96"],
                                                    ["2-hip", "2", "640"]]
    assert rows[3] == "2-hip,2,640,3,0,1.000000,2.000000,4.000000,150.0,500"


def _stub(tmp_path, body):
    p = tmp_path / "stub"
    p.write_text(f"#!{sys.executable}\nimport sys, time\n{body}\n")
    p.chmod(p.stat().st_mode | stat.S_IXUSR)
    return str(p)


def test_run_case_rows(tmp_path):
    ok = _stub(tmp_path, "print('Estimated timesteps:     ~500')\n"
                         "print('501 timesteps complete')\nprint('Time = 1.250000 seconds')")
    row, log = P.run_case(ok, "1-hip-h3d", 512, rep=2, job_id="j7")
    f = row.split(",")
    assert f[:9] == ["2", "j7", "1-hip-h3d", "1gpu", "1", "0", "512", "0.0001679668", "500"]
    assert float(f[9]) >= 0 and f[10:] == ["1.250000", "ok"]
    assert log.startswith(">>> Running: 1-hip-h3d  gpus=1  mask=0  N=512\n")
    assert P.parse_run_log(log.splitlines()) == [("1-hip-h3d", "1", "512", "1.250000", "501")]

    bad = _stub(tmp_path, "sys.exit(3)")
    assert P.run_case(bad, "1-hip-h3d", 64)[0].split(",")[10:] == ["N/A", "runtime_error"]
    slow = _stub(tmp_path, "time.sleep(30)")
    assert P.run_case(slow, "1-hip-h3d", 64, timeout=0.5)[0].split(",")[10:] == ["N/A", "timeout"]


def test_run_and_aggregate_commands(tmp_path):
    ok = _stub(tmp_path, "n = int(sys.argv[1])\nprint('%d timesteps complete' % (n + 1))\n"
                         "print('Time = %.6f seconds' % (n / 100))")
    out = tmp_path / "res"
    P.main(["run", "--bin", ok, "--sizes", "64", "128", "--reps", "2", "--out", str(out)])
    t = (out / "timings.csv").read_text().splitlines()
    assert t[0] == P.TIMINGS_HEADER and len(t) == 5
    s = (out / "summary.csv").read_text().splitlines()
    assert s == [P.SUMMARY_HEADER,
                 "1-hip-h3d,1,64,2,0,0.640000,0.640000,0.640000,0.0,65",
                 "1-hip-h3d,1,128,2,0,1.280000,1.280000,1.280000,0.0,129"]
    P.main(["aggregate", str(out / "run.log"), "-o", str(tmp_path / "s2.csv")])
    assert (tmp_path / "s2.csv").read_text().splitlines() == s


@pytest.mark.parametrize("machine", ["mi250x", "h100", "max1550"])
def test_gamma_rebuilds_reference_files(machine):
    """compute_gamma.py's gamma_single_gpu.csv for each machine, byte for
    byte, from the OpenMP summary.csv it was computed from."""
    with open(os.path.join(GOLDEN, f"gamma_{machine}_summary.csv"), newline="") as f:
        rows = P.gamma_rows(f)
    with open(os.path.join(GOLDEN, f"gamma_{machine}_single_gpu.csv"), "rb") as f:
        want = f.read()
    assert P.gamma_csv(rows).encode() == want
    assert [r["N"] for r in rows] == list(P.N_LIST)


def test_gamma_plateau_and_skip_rules():
    # the plateau is the mean of the three largest grids' recorded gammas
    with open(os.path.join(GOLDEN, "gamma_mi250x_single_gpu.csv")) as f:
        recorded = [float(line.split(",")[4]) for line in f.read().splitlines()[1:]]
    with open(os.path.join(GOLDEN, "gamma_mi250x_summary.csv"), newline="") as f:
        rows = P.gamma_rows(f)
    assert P.gamma_plateau(rows) == sum(recorded[-3:]) / 3
    assert P.gamma_plateau(rows, tail=50) == sum(recorded) / len(recorded)
    text = "\n".join([
        P.SUMMARY_HEADER,
        " 1-hip-h3d ,1,66,3,0,1.0,2.0,3.0,0.0,10",  # variant compared stripped
        "1-hip-h3d,1,2,3,0,1.0,2.0,3.0,0.0,10",    # N <= 2
        "1-hip-h3d,1,34,3,0,1.0,2.0,3.0,0.0,0",    # no steps
        "1-hip-h3d,1,34,3,0,1.0,nan,3.0,0.0,10",   # non-finite time
        "1-hip-h3d,1,34,3,0,1.0,0,3.0,0.0,10",     # zero time
        "1-hip-h3d,1,34,3,0,1.0,x,3.0,0.0,10",     # unparsable
        "1-hip,1,34,3,0,1.0,2.0,3.0,0.0,10",       # another variant
        "1-hip-h3d,1,18,3,0,1.0,0.5,3.0,0.0,4",
    ]) + "\n"
    rows = P.gamma_rows(io.StringIO(text), "1-hip-h3d")
    assert [(r["N"], r["timesteps"], r["T1_seconds"]) for r in rows] == [(18, 4, 0.5),
                                                                      (66, 10, 2.0)]
    assert rows[0]["gamma_s_per_point"] == 0.5 / (4 * 16 ** 3)
    assert P.gamma_rows(io.StringIO(text), "1-hip-h3d", "max_sec")[1]["T1_seconds"] == 3.0


def test_gamma_command(tmp_path, capsys):
    src = tmp_path / "summary.csv"
    src.write_text(P.SUMMARY_HEADER + "\n1-hip-h3d,1,512,3,0,0.16,0.164464,0.17,0.2,501\n")
    out = tmp_path / "g.csv"
    P.main(["gamma", str(src), "--exe", "1-hip-h3d", "--output", str(out)])
    assert out.read_bytes() == (
        "N,interior_points,timesteps,T1_seconds,gamma_s_per_point\r\n"
        f"512,132651000,501,0.164464,{0.164464 / (501 * 132651000)!r}\r\n").encode()
    assert "Plateau gamma (mean of N = 512)" in capsys.readouterr().out
    with pytest.raises(SystemExit):
        P.main(["gamma", str(src), "--exe", "1-omp"])
