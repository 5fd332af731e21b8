"""Writes reference_kats.json: the known-answer data of the reference's own
unit tests that pin this path, transcribed as inputs and expected outputs.

  subsample   Backend/tests/test_subsample_frame.cc:18-33, 56-69
  queue       Backend/tests/test_queue.cc:240-334 (enqueue_multiple: expected
              size and start index after each call, max_size 2)
  reduce      Backend/tests/test_reduce.cc:13-85, 91-464 (data patterns,
              sizes and the exact expected min / max / sum)
  map         Backend/tests/test_map.cc:72-98 (sizes and factor 5)
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))

img8 = list(range(1, 65))
kats = {
    "subsample": [
        {"width": 8, "height": 8, "step": 2, "input": img8,
         "expected": [1, 11, 5, 15, 18, 28, 22, 32, 33, 43, 37, 47, 50, 60, 54, 64]},
        {"width": 8, "height": 8, "step": 4, "input": img8, "expected": [1, 21, 35, 55]},
    ],
    # each case: queue max size, then a list of [n_enqueued, expected_size, expected_start]
    "queue": [
        {"name": "SimpleMultipleEnqueue", "max_size": 2, "ops": [[2, 2, 0]]},
        {"name": "CircularMultipleEnqueue", "max_size": 2, "ops": [[1, 1, 0], [2, 2, 1], [2, 2, 1], [1, 2, 0]]},
        {"name": "OversizedMultipleEnqueue", "max_size": 2, "ops": [[3, 2, 1]]},
        {"name": "FullMultipleEnqueue", "max_size": 2, "ops": [[2, 2, 0], [1, 2, 1]]},
        {"name": "DequeueMultipleFramesEnqueueOnly", "max_size": 3, "ops": [[2, 2, 0]]},
    ],
    # pattern "mod10": x[i] = i % 10 (sum); "mod1000p1": x[i] = (i % 1000) + 1 (min);
    # "neg_mod1000p1": x[i] = -((i % 1000) + 1) (max)
    "reduce": (
        [{"pattern": "mod10", "op": "sum", "width": w, "height": h}
         for (w, h) in [(1024, 1024), (1024, 512), (1025, 1025), (32, 32), (31, 31), (33, 33), (1024, 128),
                        (1024, 127), (1024, 129), (64, 64), (120, 120), (127, 1), (128, 1), (129, 1), (63, 1),
                        (64, 1), (65, 1), (29, 1), (2048, 2048), (256, 2048), (256, 1024)]]
        + [{"pattern": "mod1000p1", "op": "min", "width": 1024, "height": 1024, "expected": 1},
           {"pattern": "neg_mod1000p1", "op": "max", "width": 1024, "height": 1024, "expected": -1}]),
    "map": [{"size": 103, "value": 5.0}, {"size": 512 * 1024, "value": 5.0}],
}
for c in kats["reduce"]:
    if c["op"] == "sum":
        n = c["width"] * c["height"]
        c["expected"] = sum(i % 10 for i in range(n))

if __name__ == "__main__":
    with open(os.path.join(HERE, "reference_kats.json"), "w") as f:
        json.dump(kats, f, indent=1)
