"""Reads the reference's input fixture the way its ingest does.

construct_file_iterator_skip_newline (database/src/client.rs:163-182): every
line, split on ',', every field that `str::parse::<f64>` accepts, in order.
The README quick-start input (database/data/randomwalkdatasample1k-1k) is kept
gzip-compressed under tests/golden/ as a data fixture.
"""
import gzip
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
QUICKSTART = os.path.join(HERE, "golden", "randomwalkdatasample1k-1k.csv.gz")

# Rust's f64 grammar (no surrounding whitespace, no underscores, no hex)
_F64 = re.compile(r"^[+-]?((\d+\.?\d*|\.\d+)([eE][+-]?\d+)?|inf|infinity|nan)$", re.IGNORECASE)


def parse_fields(text):
    out = []
    for line in text.splitlines():
        for field in line.split(","):
            if _F64.match(field):
                out.append(float(field))
    return np.array(out, dtype=np.float64)


def quickstart_values():
    with gzip.open(QUICKSTART, "rt") as f:
        return parse_fields(f.read())


def quickstart_csv_text():
    with gzip.open(QUICKSTART, "rt") as f:
        return f.read()
