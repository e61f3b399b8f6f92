"""Writes hadamard_base_tables.json: the order-12 and order-20 base blocks the
reference's decoder starts from (cuda_toolkit/src/rf_processing/hadamard/
hadamard_decoder.cpp:8-45), as sign matrices.  Run once in a container that
has the reference mounted; the JSON is the committed fixture."""
import json
import re
from pathlib import Path

SRC = Path("/root/reference/cuda_toolkit/src/rf_processing/hadamard/hadamard_decoder.cpp")


def table(text, name, n):
    body = re.search(name + r"\[\]\s*=\s*\{(.*?)\};", text, re.S).group(1)
    vals = [int(v) for v in re.findall(r"-?1", body)]
    assert len(vals) == n * n, (name, len(vals))
    return [vals[i * n:(i + 1) * n] for i in range(n)]


def main():
    text = SRC.read_text()
    out = {"source": "cuda_toolkit/src/rf_processing/hadamard/hadamard_decoder.cpp:8-45",
           "h12": table(text, "hadamard_12_transpose", 12),
           "h20": table(text, "hadamard_20_transpose", 20)}
    Path(__file__).with_name("hadamard_base_tables.json").write_text(json.dumps(out))


if __name__ == "__main__":
    main()
