"""The closed-form checks have teeth: each plausible mistake, planted in a copy of the oracle's
source and compiled on the CPU, makes the check that guards it fail, while the unmodified source
passes the same checks (tests/test_oracle_closed_forms.py).

The copies are built in a temporary directory with the oracle's own compiler flags; the oracle
under oracle/ is never modified.
"""
import os
import shutil
import subprocess

import pytest

import test_oracle_closed_forms as cf
from conftest import ROOT

SRC = os.path.join(ROOT, "oracle", "wf_oracle.c")
HDR = os.path.join(ROOT, "oracle", "wf_oracle.h")

# name: (text in wf_oracle.c, replacement, the check that must fail)
MUTANTS = {
    "sun NEE without the cosine": (
        "*contrib = muls(mulv(mulv(ps->T, m->albedo), ls.Le), cs * K_INV_PI * w / ls.pdf);",
        "*contrib = muls(mulv(mulv(ps->T, m->albedo), ls.Le), K_INV_PI * w / ls.pdf);",
        lambda po, L: cf.check_sun_cosine_law(po, L)),
    "sun NEE without 1/pi": (
        "*contrib = muls(mulv(mulv(ps->T, m->albedo), ls.Le), cs * K_INV_PI * w / ls.pdf);",
        "*contrib = muls(mulv(mulv(ps->T, m->albedo), ls.Le), cs * w / ls.pdf);",
        lambda po, L: cf.check_sun_cosine_law(po, L)),
    "sun shadow ray toward -l": (
        "                wi = ls.n;\n",
        "                wi = muls(ls.n, -1.0f);\n",
        lambda po, L: cf.check_sun_occluder(po, L)),
    "sky gradient on d.x": (
        "float t = 0.5f * (d.y + 1.0f);",
        "float t = 0.5f * (d.x + 1.0f);",
        lambda po, L: cf.check_sky_direct(po, L, (1.0, -0.4, 0.3))),
    "sky gradient with lo and hi swapped": (
        "float t = 0.5f * (d.y + 1.0f);",
        "float t = 0.5f * (1.0f - d.y);",
        lambda po, L: cf.check_sky_direct(po, L, (0.0, 0.6, -1.0))),
    "escaped BSDF ray without the MIS weight": (
        "float w = power_h(ps->prev_pdf, K_INV_4PI / (float)s->n_sel);",
        "float w = 1.0f;",
        lambda po, L: cf.check_sky_nee(po, L, 8)),
    "sky NEE pdf without the selection probability": (
        "ls.pdf = K_INV_4PI / (float)s->n_sel;",
        "ls.pdf = K_INV_4PI / (float)(s->n_sel + 1);",
        lambda po, L: cf.check_sky_nee(po, L, 8)),
    "fuzz scaled by fuzz squared": (
        "if (fuzz > 0.0f) dir = add(dir, muls(sph_dir(rnd(key, depth, D_FU), rnd(key, depth, D_FV)), fuzz));",
        "if (fuzz > 0.0f) dir = add(dir, muls(sph_dir(rnd(key, depth, D_FU), rnd(key, depth, D_FV)), fuzz * fuzz));",
        lambda po, L: cf.check_fuzz(po, L, (0.0, -0.6, -0.8), 0.3)),
    "fuzz vector inside the ball instead of on the sphere": (
        "if (fuzz > 0.0f) dir = add(dir, muls(sph_dir(rnd(key, depth, D_FU), rnd(key, depth, D_FV)), fuzz));",
        "if (fuzz > 0.0f) dir = add(dir, muls(sph_dir(rnd(key, depth, D_FU), rnd(key, depth, D_FV)), "
        "fuzz * cbrtf(rnd(key, depth, 11))));",
        lambda po, L: cf.check_fuzz(po, L, (0.0, -0.6, -0.8), 0.3)),
    "fuzz from one random dimension twice": (
        "if (fuzz > 0.0f) dir = add(dir, muls(sph_dir(rnd(key, depth, D_FU), rnd(key, depth, D_FV)), fuzz));",
        "if (fuzz > 0.0f) dir = add(dir, muls(sph_dir(rnd(key, depth, D_FU), rnd(key, depth, D_FU)), fuzz));",
        lambda po, L: cf.check_fuzz(po, L, (0.0, -0.6, -0.8), 0.3)),
}


def build(tmp, name, text, repl):
    src = open(SRC).read()
    assert src.count(text) >= 1, f"mutation site of '{name}' not found in wf_oracle.c"
    d = os.path.join(tmp, "m%08x" % (abs(hash(name)) & 0xFFFFFFFF))
    os.makedirs(d, exist_ok=True)
    shutil.copy(HDR, d)
    with open(os.path.join(d, "wf_oracle.c"), "w") as f:
        f.write(src.replace(text, repl, 1))
    so = os.path.join(d, "liboracle_mut.so")
    subprocess.run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-o", so,
                    os.path.join(d, "wf_oracle.c"), "-lm", "-lpthread"], check=True)
    return so


@pytest.fixture(scope="module")
def tmpdir_m(tmp_path_factory):
    return str(tmp_path_factory.mktemp("oracle_mutants"))


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_planted_mistake_fails_its_closed_form(po, tmpdir_m, name):
    text, repl, check = MUTANTS[name]
    lib = po.load(build(tmpdir_m, name, text, repl))
    with pytest.raises(AssertionError):
        check(po, lib)


def test_unmodified_copy_passes_every_check(po, tmpdir_m):
    lib = po.load(build(tmpdir_m, "none", "#include", "#include"))
    for _, _, check in MUTANTS.values():
        check(po, lib)
