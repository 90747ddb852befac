"""The known-answer tests pin the oracle: each single mistake planted in a copy of its C source
(a dropped term, a flipped sign, a swapped operand, a wrong constant, a reference quirk undone)
makes tests/test_oracle_kats.py (or the schedule table of tests/test_schedule_and_synth.py) fail.  Each mutant is built with gcc into a temporary directory
and the KAT module runs against it in a child process (PDF_ORACLE_LIB); the unmutated source
built the same way must pass.  CPU only; about a minute on 8 cores."""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "pdflow_oracle.c")

# name -> (text of the oracle, its replacement); each text occurs once in the source
MUTATIONS = {
    "pyramid depth threshold": ("const float max_depth_dif = 0.1f;", "const float max_depth_dif = 0.2f;"),
    "pyramid depth weight without the distance term":
        ("const float w = g * (max_depth_dif - fabsf(d - dcen));", "const float w = g;"),
    "pyramid border colour not normalised": ("col[index] = inner ? sumc : sumc / wc;", "col[index] = sumc;"),
    "back-projection offset sign":
        ("((float)u - disp_u) * dep[index] * inv_f_i", "((float)u + disp_u) * dep[index] * inv_f_i"),
    "upsampled flow not doubled": ("c->f[UPS[k]][big] = (k < 2) ? 2.f * x : x;", "c->f[UPS[k]][big] = x;"),
    "upsampling border weight": ("aw -= 0.25f * g;", "aw -= 0.5f * g;"),
    "rij without the factor 2":
        ("F(PDF_RI)[index] = (fabsf(dxu) + fabsf(dzu) > 0.f) ? 2.f * rhyp(dxu, dzu) : 1.f;",
         "F(PDF_RI)[index] = (fabsf(dxu) + fabsf(dzu) > 0.f) ? rhyp(dxu, dzu) : 1.f;"),
    "rij of frame 2 from frame 1":
        ("dxu2 = xn[index + rows_i] - xn[index];\n        dzu2 = zn[index + rows_i] - zn[index];",
         "dxu2 = xo[index + rows_i] - xo[index];\n        dzu2 = zo[index + rows_i] - zo[index];"),
    "gradient weights swapped":
        ("F(PDF_DCU_AUX)[index] = (ri2[index] * (I[index + r] - I[index]) +\n                                 ri2[index - r] * (I[index] - I[index - r]))",
         "F(PDF_DCU_AUX)[index] = (ri2[index - r] * (I[index + r] - I[index]) +\n                                 ri2[index] * (I[index] - I[index - r]))"),
    "gradient weighted by frame 1's rij":
        ("const float *ri2 = F(PDF_RI_2), *rj2 = F(PDF_RJ_2);", "const float *ri2 = F(PDF_RI), *rj2 = F(PDF_RJ);"),
    "temporal intensity difference sign":
        ("F(PDF_DCT)[index] = wc - c->pyr[PDF_PYR_COLOUR_OLD][L][index];",
         "F(PDF_DCT)[index] = c->pyr[PDF_PYR_COLOUR_OLD][L][index] - wc;"),
    "bilinear rows swapped":
        ("return ((float)sv - iv) * bot + (iv - (float)fv) * top;", "return ((float)sv - iv) * top + (iv - (float)fv) * bot;"),
    "depth hole fallback returns zero": ("return mat[rv + rows_i * ru];", "return 0.f;"),
    "mu coefficient": ("1.f + 1000.f * (ddu * ddu", "1.f + 100.f * (ddu * ddu"),
    "sigma_pd without ddv":
        ("1.f / (mu_uv * (1.f + fabsf(ddu) + fabsf(ddv)) + 1e-10f)", "1.f / (mu_uv * (1.f + fabsf(ddu)) + 1e-10f)"),
    "tau without the left neighbour's ri":
        ("if (u > 0) acu_r += F(PDF_RI)[index - rows_i];", "if (u > 1) acu_r += F(PDF_RI)[index - rows_i];"),
    "depth dual step sign":
        ("(-F(PDF_DW_ACC)[index] + F(PDF_DDT)[index]", "(F(PDF_DW_ACC)[index] + F(PDF_DDT)[index]"),
    "projection threshold": ("if (m < 1.f) { *oa = a * m; *ob = b * m; }", "if (m < 0.5f) { *oa = a * m; *ob = b * m; }"),
    "divergence last column": ("if (last) return -r[index - stride] * p[index - stride];", "if (last) return 0.f;"),
    "prox case boundary": ("if (of < -th) {", "if (of < th) {"),
    "no over-relaxation": ("F(PDF_DU_ACC)[index] = 2.f * du - du_old;", "F(PDF_DU_ACC)[index] = du;"),
    "saturation bound": ("if (du > 1.f) du = 1.f; else if (du < -1.f) du = -1.f;",
                         "if (du > 2.f) du = 2.f; else if (du < -2.f) du = -2.f;"),
    "median weight of the depth difference": ("pr[n] = 1.f / (1.f + 5.f * (dz * dz)", "pr[n] = 1.f / (1.f + 10.f * (dz * dz)"),
    "median weight of the intensity change": ("+ 10.f * (dt * dt));", "+ 5.f * (dt * dt));"),
    "warped depth difference kept at holes":
        ("F(PDF_DDT)[index] = (wd > 0.f) ? wd - c->pyr[PDF_PYR_DEPTH_OLD][L][index] : 0.f;",
         "F(PDF_DDT)[index] = wd - c->pyr[PDF_PYR_DEPTH_OLD][L][index];"),
    "warp clamp at the last column":
        ("else if (*iu > (float)cols_i - 1.f) *iu = (float)cols_i - 1.f;",
         "else if (*iu > (float)cols_i - 2.f) *iu = (float)cols_i - 2.f;"),
    "schedule step": ("out[i] = out[i + 1] - 15;", "out[i] = out[i + 1] - 10;"),
    "median interpolation ends swapped":
        ("return ((cum[r] - med) * fld[l] + (med - cum[l]) * fld[r]) / (cum[r] - cum[l]);",
         "return ((cum[r] - med) * fld[r] + (med - cum[l]) * fld[l]) / (cum[r] - cum[l]);"),
    "motion field axes": ("F(PDF_DX)[index] = dwl;", "F(PDF_DX)[index] = dul;"),
    "motion field depth term": ("+ dwl * c->pyr[PDF_PYR_XX_OLD][L][index] / z;", "+ dwl * c->pyr[PDF_PYR_YY_OLD][L][index] / z;"),
}

CFLAGS = ["-O1", "-fPIC", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-shared"]


def _build(src_text, out_dir):
    c = os.path.join(out_dir, "pdflow_oracle.c")
    with open(c, "w") as f:
        f.write(src_text)
    so = os.path.join(out_dir, "liboracle.so")
    subprocess.run(["gcc", *CFLAGS, "-I", os.path.join(ROOT, "oracle"), "-o", so, c, "-lm"], check=True,
                   capture_output=True)
    return so


def _kats(so):
    env = dict(os.environ, PDF_ORACLE_LIB=so, OMP_NUM_THREADS="1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_oracle_kats.py"),
                        os.path.join(ROOT, "tests", "test_schedule_and_synth.py") + "::test_schedule_table"],
                       capture_output=True, text=True, env=env, cwd=ROOT, timeout=600)
    return r.returncode, (r.stdout.strip().splitlines() or [""])[-1]


@pytest.fixture(scope="module")
def src():
    if not shutil.which("gcc"):
        pytest.skip("gcc is not available")
    with open(SRC) as f:
        return f.read()


def test_mutations_apply_once(src):
    for name, (a, _) in MUTATIONS.items():
        assert src.count(a) == 1, name


def test_every_planted_mistake_fails_the_known_answer_tests(src, tmp_path):
    def run(item):
        name, text = item
        d = tmp_path / str(abs(hash(name)))
        d.mkdir()
        return name, _kats(_build(text, str(d)))

    jobs = [("(unmutated)", src)] + [(n, src.replace(a, b)) for n, (a, b) in MUTATIONS.items()]
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        results = dict(ex.map(run, jobs))
    rc, last = results.pop("(unmutated)")
    assert rc == 0, f"the unmutated oracle fails its known-answer tests: {last}"
    survivors = [f"{n}: {last}" for n, (rc, last) in results.items() if rc != 1]
    assert not survivors, "mistakes the known-answer tests do not catch:\n" + "\n".join(survivors)
