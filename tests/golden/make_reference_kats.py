"""Writes reference_kats.json: the known-answer data of the reference's own
unit tests that pin this path, transcribed as inputs and expected outputs.

  subsample   Backend/tests/test_subsample_frame.cc:18-33, 56-69
  queue       Backend/tests/test_queue.cc:240-334 (enqueue_multiple: expected
              size and start index after each call, max_size 2)
  reduce      Backend/tests/test_reduce.cc:13-85, 91-464 (data patterns,
              sizes and the exact expected min / max / sum)
  map         Backend/tests/test_map.cc:72-98 (sizes and factor 5)
  batch_input_queue
              Backend/tests/test_batch_input_queue.cc:48-287: frame capacity,
              batch size, frame width (height 1, 8-bit), the frames (C strings
              padded with NUL to the frame size), then the operations with
              the size expected after each: ["enq", frame], ["deq", expected
              frames or null when the test does not check them], ["resize",
              new batch size], ["elt", position, expected frame]; the size
              is null where the test does not check it
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
    "batch_input_queue": [
        {"name": "SimpleInstantiation", "capacity": 3, "batch": 1, "width": 2, "height": 2, "frames": [],
         "ops": [["size", 0]]},
        {"name": "SimpleEnqueueOfThreeElements", "capacity": 3, "batch": 1, "width": 2, "height": 1,
         "frames": ["a", "b", "c", "d", "e"],
         "ops": [["enq", 0, None], ["enq", 1, None], ["enq", 2, 3], ["elt", 0, "a", 3], ["elt", 1, "b", 3],
                 ["elt", 2, "c", 3]]},
        {"name": "SimpleEnqueueAndDequeueOfThreeElements", "capacity": 3, "batch": 1, "width": 2, "height": 1,
         "frames": ["a", "b", "c", "d", "e"],
         "ops": [["enq", 0, None], ["enq", 1, None], ["enq", 2, None], ["deq", ["a"], 2], ["deq", ["b"], 1],
                 ["deq", ["c"], 0]]},
        {"name": "SimpleOverwriteElements", "capacity": 3, "batch": 1, "width": 2, "height": 1,
         "frames": ["a", "b", "c", "d", "e"],
         "ops": [["enq", 0, None], ["enq", 1, None], ["enq", 2, None], ["enq", 3, None], ["enq", 4, 3],
                 ["deq", ["c"], 2],
                 ["deq", ["d"], 1], ["deq", ["e"], 0]]},
        {"name": "SimpleOverwriteMoreElements", "capacity": 4, "batch": 2, "width": 4, "height": 1,
         "frames": ["abc", "ABC", "def", "DEF", "ghi", "GHI"],
         "ops": [["enq", 0, 0], ["enq", 1, 1], ["enq", 2, 1], ["enq", 3, 2], ["deq", ["abc", "ABC"], 1],
                 ["enq", 4, 1], ["enq", 5, 2], ["enq", 0, 2], ["enq", 1, 2], ["deq", ["ghi", "GHI"], 1],
                 ["deq", ["abc", "ABC"], 0]]},
        {"name": "SimpleResizeSame", "capacity": 4, "batch": 2, "width": 5, "height": 1,
         "frames": ["ilan", "nico", "anto", "kaci", "theo"],
         "ops": [["enq", 0, None], ["enq", 1, None], ["enq", 2, None], ["enq", 3, None], ["deq", None, 1],
                 ["resize", 2, 0], ["enq", 4, None], ["enq", 0, 1], ["deq", None, 0]]},
        {"name": "SimpleResizeGreater", "capacity": 4, "batch": 2, "width": 5, "height": 1,
         "frames": ["ilan", "nico", "anto", "kaci", "theo"],
         "ops": [["enq", 0, None], ["enq", 1, None], ["enq", 2, None], ["enq", 3, None], ["deq", None, 1],
                 ["resize", 4, 0], ["enq", 4, None], ["enq", 0, None], ["enq", 2, None], ["enq", 3, 1],
                 ["deq", None, 0]]},
        {"name": "SimpleResizeLower", "capacity": 4, "batch": 2, "width": 5, "height": 1,
         "frames": ["ilan", "nico", "anto", "kaci", "theo"],
         "ops": [["enq", 0, None], ["enq", 1, None], ["enq", 2, None], ["enq", 3, None], ["deq", None, 1],
                 ["resize", 1, 0], ["enq", 4, 1], ["deq", None, 0]]},
    ],
}
for c in kats["reduce"]:
    if c["op"] == "sum":
        n = c["width"] * c["height"]
        c["expected"] = sum(i % 10 for i in range(n))

if __name__ == "__main__":
    with open(os.path.join(HERE, "reference_kats.json"), "w") as f:
        json.dump(kats, f, indent=1)
