"""Writes known_shapes.json: the polyomino tables the reference ships
(src/known1.cpp:3-14, free; src/known2.cpp:3-14, one-sided) as data.

Run in a container that has the reference at /root/reference; the JSON is
committed so nothing reads the reference at test time."""
import json
import os
import re

REF = "/root/reference/src"
HERE = os.path.dirname(os.path.abspath(__file__))


def parse(path):
    text = open(path).read()
    arrays = {}
    for m in re.finditer(r"static const uint64_t s(\d)\[\]\{([^}]*)\}", text):
        arrays[int(m.group(1))] = [int(x) for x in m.group(2).replace(" ", "").split(",") if x]
    counts = [int(x) for x in re.search(r"shapes(?:_C)?_count\[\]\{([^}]*)\}", text)
              .group(1).replace(" ", "").split(",") if x]
    return arrays, counts


free, free_counts = parse(os.path.join(REF, "known1.cpp"))
one_sided, one_sided_counts = parse(os.path.join(REF, "known2.cpp"))
json.dump({
    "source": "src/known1.cpp:3-14 (free), src/known2.cpp:3-14 (one-sided)",
    "free": {str(k): v for k, v in sorted(free.items())},
    "free_counts": free_counts,
    "one_sided": {str(k): v for k, v in sorted(one_sided.items())},
    "one_sided_counts": one_sided_counts,
}, open(os.path.join(HERE, "known_shapes.json"), "w"), indent=0)
