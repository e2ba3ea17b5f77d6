"""Generate the model variants compiled into the HIP library.

``python -m goss_amd.codegen.generate`` (re)writes
``goss_amd/csrc/generated/<variant>.h`` (expression templates),
``<variant>_plugin.h`` (``goss::ParameterizedODE`` subclass) and
``variants.json`` (names, defaults, field selection, op census).
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
from typing import Dict, List

from .emit import Variant, emit_model_header, emit_parameterized_ode, variant_metadata
from .ode_model import load_ode

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ODE_DIR = os.path.join(PKG, "odes")
OUT_DIR = os.path.join(PKG, "csrc", "generated")

TP06_M = "tentusscher_panfilov_2006_M_cell"
TP06_EPI = "tentusscher_panfilov_2006_epi_cell"

# name -> (ode file stem, field states, field parameters)
VARIANTS: Dict[str, tuple] = {
    "tp06_m": (TP06_M, ["V"], []),
    "tp06_m_fp3": (TP06_M, ["V"], ["g_Kr", "g_Ks", "g_CaL"]),
    "tp06_m_fp6": (TP06_M, ["V"], ["g_Kr", "g_Ks", "g_CaL", "g_Na", "g_K1", "g_to"]),
    "tp06_epi": (TP06_EPI, ["V"], []),
    "fitzhughnagumo": ("fitzhughnagumo", ["V"], []),
    "lorentz": ("lorentz", ["x"], []),
    "fhn": ("fhn", ["v"], []),
    "beeler_reuter_1977": ("beeler_reuter_1977", ["V"], []),
    "oscilator": ("oscilator", ["x", "y"], ["a", "b"]),
    "linear_decay": ("linear_decay", ["y"], ["a", "b"]),
    "van_der_pol_dae": ("van_der_pol_dae", ["y"], []),
    "markov_three_state": ("markov_three_state", ["O"], ["k_oi"]),
}


def make_variant(name: str) -> Variant:
    stem, fs, fp = VARIANTS[name]
    model = load_ode(os.path.join(ODE_DIR, stem + ".ode"))
    return Variant(model=model, name=name, field_states=list(fs), field_parameters=list(fp))


def generate(names: List[str] = None, out_dir: str = OUT_DIR) -> Dict[str, dict]:
    names = names or list(VARIANTS)
    os.makedirs(out_dir, exist_ok=True)
    meta: Dict[str, dict] = {}
    for n in names:
        var = make_variant(n)
        em = emit_model_header(var)
        _write(os.path.join(out_dir, f"{n}.h"), em.header)
        _write(os.path.join(out_dir, f"{n}_plugin.h"), emit_parameterized_ode(var, f"{n}.h"))
        meta[n] = variant_metadata(var, em)
        stem = VARIANTS[n][0]
        with open(os.path.join(ODE_DIR, stem + ".ode"), "rb") as f:
            meta[n]["source_sha256"] = hashlib.sha256(f.read()).hexdigest()
    _write(os.path.join(out_dir, "variants.json"), json.dumps(meta, indent=1, sort_keys=True) + "\n")
    lines = ["// Generated by goss_amd.codegen.generate -- do not edit.",
             "// X-macro list of the model variants compiled into the library.",
             "#pragma once",
             "#define GOSS_VARIANT_LIST(X) \\"]
    lines += [f"  X({n}) \\" for n in names]
    lines += ["", ""]
    _write(os.path.join(out_dir, "variant_list.h"), "\n".join(lines))
    return meta


def _write(path: str, text: str) -> None:
    # Rewrite only on change so that make-style rebuilds stay incremental.
    if os.path.exists(path):
        with open(path) as f:
            if f.read() == text:
                return
    with open(path, "w") as f:
        f.write(text)


if __name__ == "__main__":
    generate(sys.argv[1:] or None)
