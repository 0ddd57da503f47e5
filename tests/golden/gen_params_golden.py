"""Write tests/golden/params_golden.json: parameter-file cases run through the
reference's own reader (read_param.c + confirm_param.c + msg.c, built from the
reference tree by `make -C oracle -f ref.mk` into oracle/_ref/ref_params).

Each case records the input text, whether confirm_parameters ran, and either
the parsed values or the abort code the reference ended with (MPI_Abort).

Usage: python tests/golden/gen_params_golden.py
"""
import json
import os
import re
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "ref_params")

BASE = """nc = 64
boxsize = 100.0
random_seed= 100
ntimestep= 10
a_final= 1.0
output_redshifts= {0.73, 0.6, 0.44, 0.0}
omega_m = 0.273
h       = 0.705
sigma8  = 0.812
powerspectrum= "data/linear_matterpower.dat"
"""

CASES = {
    "full_param_file": (open(os.path.join(HERE, "param_full.lua")).read(), 1),
    "minimal": (BASE, 1),
    "no_spaces": ("nc=32\nboxsize=50\nntimestep=5\noutput_redshifts={0}\nomega_m=0.3\nh=0.7\nsigma8=0.8\n"
                  "powerspectrum=\"pk.dat\"\n", 1),
    "double_equals": ("nc == 16\nboxsize ===  25.5\n", 0),
    "leading_blanks_and_tabs": ("   \tnc = 8\n\t\tboxsize =\t12.5\n", 0),
    "tab_before_equals_is_part_of_name": ("nc\t= 8\n", 0),
    "negative_number_reads_zero": ("nc = -64\nboxsize = -1.0\nrandom_seed = +7\na_final = .5\n", 0),
    "exponent_and_trailing_text": ("boxsize = 1e3 Mpc\nomega_m = 0.3abc\nnc = 12.9\n", 0),
    "comment_inside_string": ("powerspectrum = \"a--b.dat\"\n", 0),
    "comment_lines": ("-- header\n   -- indented comment\n\nnc = 4 -- four\n", 0),
    "string_unterminated": ("snapshot = \"snp\n", 0),
    "string_empty": ("fof = \"\"\n", 0),
    "string_without_quote": ("fof = fof\n", 0),
    "string_with_spaces": ("snapshot = \"my snap\" trailing\n", 0),
    "array_blank_before_comma": ("output_redshifts = {1 , 2}\n", 0),
    "array_blanks_after_comma": ("output_redshifts = {  3,   2.5,1}\n", 0),
    "array_signs": ("output_redshifts = {-0.5, +1, .25, 2e-1}\n", 0),
    "array_empty": ("output_redshifts = {}\n", 0),
    "array_unclosed": ("output_redshifts = {1, 2\n", 0),
    "array_no_brace": ("output_redshifts = 1, 2\n", 0),
    "array_bad_token": ("output_redshifts = {1, x}\n", 0),
    "array_trailing_comma": ("output_redshifts = {1, 2,}\n", 0),
    "array_many": ("output_redshifts = {" + ", ".join(str(i / 10) for i in range(40)) + "}\n", 0),
    "bool_values": ("write_longid = T\n", 0),
    "bool_true_word": ("write_longid = true\n", 0),
    "bool_one": ("write_longid = 1\n", 0),
    "bool_false_word": ("write_longid = false\n", 0),
    "bool_bad": ("write_longid = yes\n", 0),
    "unknown_key": ("nc = 4\nfoo = 3\n", 0),
    "duplicate_key_last_wins": ("nc = 4\nnc = 8\nfof = \"a\"\nfof = \"bb\"\n", 0),
    "crlf_line_ends": ("nc = 16\r\nboxsize = 20\r\nfof = \"f\"\r\n", 0),
    "all_options": (BASE + "pm_nc_factor = 2\nnp_alloc_factor = 1.5\nloglevel = 1\nnrealization = 3\n"
                    "fof = \"fof\"\nlinking_factor = 0.25\nsnapshot = \"snp\"\nsubsample = \"sub\"\n"
                    "subsample_factor = 0.02\ncoarse_grid = \"grid\"\ncoarse_grid_nc = 32\n"
                    "initial = \"init\"\nwrite_longid = true\n", 1),
    "confirm_missing_nc": (BASE.replace("nc = 64\n", ""), 1),
    "confirm_missing_box": (BASE.replace("boxsize = 100.0\n", ""), 1),
    "confirm_missing_steps": (BASE.replace("ntimestep= 10\n", ""), 1),
    "confirm_missing_zout": (BASE.replace("output_redshifts= {0.73, 0.6, 0.44, 0.0}\n", ""), 1),
    "confirm_missing_omega": (BASE.replace("omega_m = 0.273\n", ""), 1),
    "confirm_missing_h": (BASE.replace("h       = 0.705\n", ""), 1),
    "confirm_missing_sigma8": (BASE.replace("sigma8  = 0.812\n", ""), 1),
    "confirm_missing_power": (BASE.replace("powerspectrum= \"data/linear_matterpower.dat\"\n", ""), 1),
    "confirm_empty_power": (BASE.replace("\"data/linear_matterpower.dat\"", "\"\""), 1),
    "confirm_empty_fof": (BASE + "fof = \"\"\n", 1),
    "confirm_empty_snapshot": (BASE + "snapshot = \"\"\n", 1),
    "confirm_empty_subsample": (BASE + "subsample = \"\"\n", 1),
    "confirm_empty_grid": (BASE + "coarse_grid = \"\"\n", 1),
    "confirm_empty_initial": (BASE + "initial = \"\"\n", 1),
    "confirm_zero_sigma8_literal": (BASE.replace("sigma8  = 0.812", "sigma8 = -0.8"), 1),
}


def parse_output(text):
    out = {"zout": []}
    body = text.split("--- parameters\n", 1)[1]
    for line in body.splitlines():
        key, val = line.split("=", 1)
        if key == "zout":
            out["zout"].append(float(val))
        elif key.endswith("_filename"):
            m = re.match(r"^(<absent>|\[(.*)\]) (\d+)$", val, re.S)
            out[key] = None if m.group(1) == "<absent>" else m.group(2)
            out["strlen_" + key] = int(m.group(3))
        elif key in ("np_alloc_factor", "a_final", "boxsize", "omega_m", "sigma8", "h", "fof_linking_factor",
                     "subsample_factor"):
            out[key] = float(val)
        else:
            out[key] = int(val)
    return out


def run_reference(text, confirm):
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "param.lua")
        with open(path, "w", newline="") as f:
            f.write(text)
        r = subprocess.run([REF_BIN, str(confirm), path], capture_output=True, text=True, timeout=30)
    # the exit status holds only the low 8 bits of the abort code (11008 -> 0)
    m = re.search(r"MPI_Abort\(MPI_COMM_WORLD, (\d+)\)", r.stderr)
    if m:
        return {"rc": int(m.group(1)), "params": None}
    assert r.returncode == 0, r.stderr
    return {"rc": 0, "params": parse_output(r.stdout)}


def main():
    cases = {}
    for name, (text, confirm) in CASES.items():
        cases[name] = {"text": text, "confirm": confirm, **run_reference(text, confirm)}
    with open(os.path.join(HERE, "params_golden.json"), "w") as f:
        json.dump({"source": "reference read_param.c / confirm_param.c / msg.c via oracle/_ref/ref_params",
                   "cases": cases}, f, indent=1)


if __name__ == "__main__":
    main()
