"""bench.py's command-line contract, checked without a GPU."""
import json
import sys
from pathlib import Path

sys.path.insert(0, str(Path(__file__).resolve().parents[1]))

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_one_json_line(capsys):
    rc = bench.main(["--gpus", "2"])
    out = capsys.readouterr().out.strip().splitlines()
    assert rc == 2
    assert len(out) == 1
    line = json.loads(out[0])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]


def test_every_config_names_its_workload():
    for name, cfg in bench.CONFIGS.items():
        assert cfg["workload"], name
        if "gen" in cfg:
            assert cfg["k"] >= 1 and cfg["rounds"] >= 1 and cfg["eps"] > 0, name
