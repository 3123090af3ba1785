"""Extract the reference's own test data into JSON fixtures.

Run in a container where the reference tree is mounted; the output
(tests/golden/fixtures.json) is committed, so nothing at test time reads the
reference.  Only data is taken: the FASTA inputs the reference's tests use
(data/*.fasta), the PSSM text files of data/ its matrix reader parses, the
known-answer values of tests/test_basic.c, and the two substitution matrices
those known answers need (blosum62 and pam50 value tables, parsed as numbers
from parasail/matrices/*.h).

    python tests/golden/make_fixtures.py /root/reference
"""
import json
import os
import re
import sys


def read_fasta(path):
    recs, name, seq = [], None, []
    with open(path) as fh:
        for line in fh:
            line = line.rstrip("\n\r")
            if line.startswith(">"):
                if name is not None:
                    recs.append((name, "".join(seq)))
                name, seq = line[1:].strip(), []
            elif line:
                seq.append(line.strip())
    if name is not None:
        recs.append((name, "".join(seq)))
    return recs


def read_matrix_header(path):
    text = open(path).read()
    body = text[text.index("_[] = {") + 7:]
    body = body[:body.index("};")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    vals = [int(v) for v in re.findall(r"-?\d+", body)]
    tail = text[text.index("parasail_matrix_t"):]
    alpha = re.search(r'"([A-Z*]+)"', tail[tail.index("TYPE_SQUARE"):]).group(1)
    nums = re.findall(r"^\s*(-?\d+),\s*$", tail, flags=re.M)
    size, mx, mn = int(nums[0]), int(nums[1]), int(nums[2])
    assert len(vals) == size * size
    return {"alphabet": alpha, "size": size, "max": mx, "min": mn, "values": vals}


def main(ref):
    out = {"fasta": {}, "matrices": {}, "kat": [], "matrix_files": {}}
    for fn in ["test_small_2.fasta", "test_P0_2.fasta", "test_P0_3.fasta",
               "test_P0_4.fasta", "test_P0_5.fasta", "test_P0_10.fasta",
               "test_case.fasta", "test_case2.fasta", "test_64.fasta",
               "test_64r.fasta", "test_p29.fasta"]:
        out["fasta"][fn] = read_fasta(os.path.join(ref, "data", fn))
    # PSSM text files the reference's matrix reader takes (data/*.txt,
    # read by src/io.c:921-1178 through tests/test_matrix.c), kept as text
    out["matrix_files"] = {}
    for fn in ["cd02249_matrix_safe.txt", "cd02249_matrix_noquery.txt"]:
        with open(os.path.join(ref, "data", fn)) as fh:
            out["matrix_files"][fn] = fh.read()
    for name in ["blosum62", "pam50"]:
        out["matrices"][name] = read_matrix_header(
            os.path.join(ref, "parasail", "matrices", name + ".h"))
    # tests/test_basic.c:20-85
    out["kat"] = [
        {"fn": "sw", "s1": "asdf", "s2": "asdf", "open": 10, "gap": 1,
         "matrix": "blosum62", "score": 20, "src": "tests/test_basic.c:22-30"},
        {"fn": "sw", "s1": "asdf", "s2": "asdf", "open": 10, "gap": 1,
         "matrix": "pam50", "score": 27, "src": "tests/test_basic.c:34-42"},
        {"fn": "sw", "s1": "acgt", "s2": "acgt", "open": 10, "gap": 1,
         "matrix": ["acgt", 1, -1], "score": 4, "src": "tests/test_basic.c:46-56"},
        {"fn": "sw_striped_profile_8", "s1": "asdf", "s2": "asdf", "open": 10,
         "gap": 1, "matrix": "blosum62", "score": 20,
         "src": "tests/test_basic.c:62-83"},
    ]
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures.json")
    with open(dst, "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
