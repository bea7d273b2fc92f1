"""Writes example2_schedules.json: the two asynchronous-iteration schedules
the reference printed for its example2 DAG (10 clusters; 1 and 10
iterations; every task 0.1 on both processors, 0.1 to cross), pasted into
documents/example2_scheduling_result.py:309-322 as the four lists that
Scheduler::print_schedule_to_graph_xpbd prints:

  cpu, gpu  (colour, task, buffer, start, end) per scheduled task
  ins       (proc, task, proc, task, end, start) per cross-processor input
  left      (proc, task, proc, task, end, start) per same-processor "left" link

Run where the reference checkout is present (tests read only the fixtures):
    python tests/golden/make_example2_schedules.py
"""
import ast
import json
import os
import re

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def parse_dump(text, label):
    start = text.index(f"visualize_schedule3([ # {label}\n")
    body = text[start:text.index("], vbd_task_name_map", start)]
    lists = re.findall(r"^\s*\[(.*)\],\s*$", body, re.M)
    assert len(lists) == 4, (label, len(lists))
    return [[list(t) for t in ast.literal_eval("[" + l + "]")] for l in lists]


def main():
    text = open(os.path.join(REF, "documents/example2_scheduling_result.py")).read()
    out = {"source": "documents/example2_scheduling_result.py:309-322"}
    for label, iters in (("Iter 1", 1), ("Iter 10", 10)):
        cpu, gpu, ins, left = parse_dump(text, label)
        out[f"iters{iters}"] = {"clusters": 10, "cpu": cpu, "gpu": gpu, "ins": ins, "left": left}
    with open(os.path.join(HERE, "example2_schedules.json"), "w") as f:
        json.dump(out, f)
    print({k: {kk: len(vv) for kk, vv in v.items() if isinstance(vv, list)} for k, v in out.items() if k != "source"})


if __name__ == "__main__":
    main()
