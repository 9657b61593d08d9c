"""Writes excitation_references.json: the reference levels of the excitation solver recorded in
the reference's test/fixtures/excitation_references.toml (states, settings, and per level the
Ritz energy, the four validation energies, the validated energy and the tolerances), as data.

Run from the repository root: python tests/golden/make_excitation_references.py
"""
import json
from pathlib import Path

import tomli

SRC = Path("/root/reference/test/fixtures/excitation_references.toml")
KEEP = ("name", "state", "settings", "sector", "momentum", "nev", "level", "ritz_energy", "ritz_tolerance",
        "four_energies", "validated_energy", "tolerance", "residual", "numops", "start", "accepted",
        "converged", "dimension")


def main():
    table = tomli.loads(SRC.read_text())
    out = {"source": "test/fixtures/excitation_references.toml (format_version %d)" % table["format_version"],
           "convention": table["convention"],
           "states": table["states"], "settings": table["settings"],
           "levels": [{k: e[k] for k in KEEP if k in e} for e in table["levels"]]}
    path = Path(__file__).with_name("excitation_references.json")
    path.write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {path} ({len(out['levels'])} levels)")


if __name__ == "__main__":
    main()
