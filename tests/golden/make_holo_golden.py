"""Writes the .holo fixtures from the reference's test data (run in a tree
that has the reference source at /root/reference; the outputs are committed).

  holo_recordings.json   name and byte size of each recording the reference's
                         tests read (tests/data/inputs/*.holo, tests/*.holo).
                         The recordings are Git LFS objects that are not in
                         the source tree: only their pointer files (which
                         state the size) are, so the sizes are the data kept.
  holo_settings/*.json   the compute-settings files of the reference's CLI
                         test cases (tests/data/<case>/holovibes.json,
                         version "v1"), copied as they are, for the cases
                         whose names state the settings they exercise.
"""
import json
import os
import re
import shutil

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/tests"

CASES = [
    "8frames_64_64_space_none_time_none_no_args", "8frames_64_64_space_none_time_stft_no_args",
    "8frames_64_64_space_none_time_pca_no_args", "8frames_64_64_space_none_time_ssatft_no_args",
    "8frames_64_64_space_1fft_time_none_no_args", "8frames_64_64_space_1fft_time_stft_no_args",
    "8frames_64_64_space_1fft_time_pca_no_args", "8frames_64_64_space_1fft_time_ssatft_no_args",
    "8frames_64_64_space_2fft_time_none_no_args", "8frames_64_64_space_2fft_time_stft__no_args",
    "8frames_64_64_space_2fft_time_pca_no_args", "8frames_64_64_space_2fft_time_ssatft_no_args",
    "8frames_1024_1024_raw", "8frames_1024_1024_time_stride_3", "8frames_1024_1024_batch_size_30",
    "8frames_64_64_g",
]


def main():
    recs = []
    for d in (os.path.join(REF, "data", "inputs"), REF):
        for f in sorted(os.listdir(d)):
            if not f.endswith(".holo"):
                continue
            text = open(os.path.join(d, f), errors="replace").read(400)
            m = re.search(r"^size (\d+)$", text, re.M)
            if m and text.startswith("version https://git-lfs"):
                recs.append({"name": f, "size": int(m.group(1))})
    with open(os.path.join(HERE, "holo_recordings.json"), "w") as fh:
        json.dump(recs, fh, indent=1)
    out = os.path.join(HERE, "holo_settings")
    os.makedirs(out, exist_ok=True)
    for c in CASES:
        shutil.copyfile(os.path.join(REF, "data", c, "holovibes.json"), os.path.join(out, c + ".json"))
    print(f"{len(recs)} recordings, {len(CASES)} settings files")


if __name__ == "__main__":
    main()
