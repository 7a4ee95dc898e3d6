"""bench.py's contract on the CPU: a request for more than one GPU exits with
a message before any GPU call, and the committed traffic figures
(profiles/traffic.json) cover the plans the headline form can pick, under the
plan descriptions the planner produces now, so the benchmark line carries
roofline.traffic whichever of them the autotuner keeps."""
import os
import subprocess
import sys

from tccg_amd import Contraction
from tccg_amd.plan import F32, build_packed_plan
from tccg_amd.workloads import get

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_headline_plans_have_profiled_traffic():
    import bench
    c = Contraction.from_tccg(get("abcd-aebf-dfce"))
    for tile in (3, 7):     # 192x192x16 and 192x192x32, stream-k over 256 workgroups
        plan = build_packed_plan(c.problem, F32, tile, stream_k=1)
        traffic = bench.profiled_traffic(str(c), "f32", plan.describe())
        assert traffic is not None and traffic > 0, plan.describe()
