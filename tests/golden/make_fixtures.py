"""Copies the reference's own golden data into tests/golden/ (run where /root/reference exists).

Fixtures written:
  example_center/*.csv, parameters.txt   the small files of mpi2d/example_center (a shipped
                                          500-step 2D run), verbatim
  example_center/sha256.json              SHA-256 of all 7 files (mpi2d/MANIFEST.sha256:6-12)
  example_center/args.json                the command line that produced them
                                          (mpi2d/example_center/parameters.txt:2-64)
  blocks_regression.json                  SHA-256 pins of the 10 mpi3d-blocks regression
                                          cases at one rank (tests/logs/output_regression_summary.json)
                                          with their arguments (tests/verify_mpi.py:119-133)
These are data only: inputs and expected outputs.
"""
import json
import os
import shutil

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    ex = os.path.join(HERE, "example_center")
    os.makedirs(ex, exist_ok=True)
    src = os.path.join(REF, "mpi2d", "example_center")
    for f in ("diagnostics.csv", "grain_diagnostics.csv", "grain_history.csv", "grains.csv", "parameters.txt"):
        shutil.copyfile(os.path.join(src, f), os.path.join(ex, f))
    sha = {}
    with open(os.path.join(REF, "mpi2d", "MANIFEST.sha256")) as fh:
        for line in fh:
            h, name = line.split()
            if name.startswith("example_center/"):
                sha[name.split("/", 1)[1]] = h
    json.dump(sha, open(os.path.join(ex, "sha256.json"), "w"), indent=1, sort_keys=True)
    args = ["--nx", "61", "--ny", "61", "--grain-diameter-grid", "50", "--seed", "2001", "--left-edge-sites", "2",
            "--grain-evolution", "on", "--eta-output", "all", "--nucleation", "center", "--radius-grid", "5.5",
            "--bc", "closed", "--ck", "off", "--bounds", "legacy", "--dt", "1e-12", "--steps", "500",
            "--output-every", "500"]
    json.dump({"args": args}, open(os.path.join(ex, "args.json"), "w"), indent=1)

    summary = json.load(open(os.path.join(REF, "mpi3d-blocks", "tests", "logs", "output_regression_summary.json")))
    cases = {
        'fixed_mu': ['--grain-evolution', 'off', '--bc', 'mu', '--nucleation', 'off'],
        'eta_mu': ['--grain-evolution', 'on', '--bc', 'mu', '--nucleation', 'off'],
        'anchored': ['--grain-evolution', 'on', '--bc', 'mu', '--left-face-sites', '2', '--nucleation', 'off'],
        'center': ['--grain-evolution', 'on', '--bc', 'closed', '--nucleation', 'center'],
        'fixed_center': ['--grain-evolution', 'off', '--bc', 'closed', '--nucleation', 'center'],
        'O_off': ['--grain-evolution', 'on', '--bc', 'closed', '--nucleation', 'center', '--oxygen-diffusion', 'off'],
        'legacy_Ck': ['--grain-evolution', 'on', '--bc', 'closed', '--nucleation', 'center', '--ck', 'legacy'],
        'eta_frozen': ['--grain-evolution', 'on', '--eta-L', '0', '--bc', 'mu', '--nucleation', 'off'],
        'eta_concentration': ['--grain-evolution', 'on', '--bc', 'concentration', '--nucleation', 'off'],
        'single_crystal': ['--grains', 'off', '--bc', 'concentration', '--nucleation', 'off'],
    }
    common = ['--nx', '31', '--ny', '17', '--nz', '15', '--grain-diameter-grid', '18', '--radius-grid', '3.5',
              '--seed', '2001', '--eta-output', 'all', '--ck', 'off', '--bounds', 'legacy', '--dt', '1e-12',
              '--steps', '30', '--output-every', '30']
    out = {"common_args": common, "cases": {}}
    for rec in summary:
        if rec.get("layout") == "auto_np1":
            out["cases"][rec["case"]] = {"args": cases[rec["case"]], "sha256": rec["sha256"]}
    json.dump(out, open(os.path.join(HERE, "blocks_regression.json"), "w"), indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
