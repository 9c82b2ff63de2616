"""bench.py pieces that run without a GPU: the traffic figure is read from
the committed rocprof summary only for the SpMV layout it was measured on,
and a multi-GPU request exits with a message (one GPU is measured)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def _profile_lines():
    out = {}
    with open(bench.TRAFFIC_PROFILE) as f:
        for line in f:
            w = line.split()
            if len(w) in (7, 9) and w[0] == "traffic_bytes_per_application":
                layout = tuple(int(x) for x in w[3:]) + (0, 0)[: 9 - len(w)]
                out[(w[1], layout)] = float(w[2])
    return out


def test_traffic_lookup_matches_profile():
    lines = _profile_lines()
    assert lines, "the traffic summary holds machine-readable lines"
    assert any(layout[4:] == (0, 0) for _, layout in lines)      # row-block layout
    assert any(layout[4:] != (0, 0) for _, layout in lines)      # column-sorted layout
    for (kclass, layout), value in lines.items():
        assert bench.measured_traffic(kclass, layout[:2], layout[2:4], layout[4:]) == value
        if layout[4:] == (0, 0):
            assert bench.measured_traffic(kclass, layout[:2], layout[2:4]) == value
        # more than algorithmic bytes (about 0.42 GB per application), less than 3x
        assert 4.0e8 < value < 1.3e9
    assert bench.measured_traffic("spmv_t", (3, 5), (0, 0)) is None
    assert bench.measured_traffic("spmv_t", (2, 4), (-1, -1), (16384, 16384)) is None
    assert bench.measured_traffic("gemvt", (4, 4), (0, -1)) is None
    # a lane shift pair that was not measured: the same panels and caps answer
    for (kclass, layout), value in lines.items():
        other = bench.measured_traffic(kclass, layout[:2], (-2, -2), layout[4:])
        assert other is not None
        same = [v for (k, l), v in lines.items() if k == kclass and l[:2] == layout[:2]
                and l[4:] == layout[4:]]
        assert other in same and abs(other - value) <= 0.02 * value


def test_multi_gpu_request_exits():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""
