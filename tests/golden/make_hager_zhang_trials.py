"""Writes hager_zhang_trials.json: the returned step and the trial steps of OptimKit 0.5.1's
HagerZhangLineSearch(epsilon=1e-10, acceptfirst=true) on ten one-dimensional functions, as the
reference's test suite records them (gpu/tests/optimizer/test_line_search.py:32-57, printed by
gpu/qualification/optimizer/hager_zhang_references.jl).  Only the literal GOLDEN table is read,
as data; the functions themselves are restated in tests/test_linesearch.py.

Run from the repository root: python tests/golden/make_hager_zhang_trials.py
"""
import ast
import json
from pathlib import Path

SRC = Path("/root/reference/gpu/tests/optimizer/test_line_search.py")


def main():
    tree = ast.parse(SRC.read_text())
    for node in tree.body:
        if isinstance(node, ast.Assign) and any(getattr(t, "id", None) == "GOLDEN" for t in node.targets):
            table = ast.literal_eval(node.value)
            break
    else:
        raise SystemExit("GOLDEN table not found")
    out = {name: {"step": step, "trials": trials} for name, (step, trials) in table.items()}
    path = Path(__file__).with_name("hager_zhang_trials.json")
    path.write_text(json.dumps({"source": "OptimKit 0.5.1 HagerZhangLineSearch(epsilon=1e-10), acceptfirst=true",
                                "cases": out}, indent=1) + "\n")
    print(f"wrote {path} ({len(out)} cases)")


if __name__ == "__main__":
    main()
