"""Mutation check of the oracle's pins (maintainer tool, CPU only).

Applies each one-line change below, one at a time, to a scratch copy of the
repository (oracle/fribo_oracle.c only), rebuilds the oracle there and runs
the CPU test suite (pytest -m "not gpu").  Every change is a plausible
mistake in restating the reference (a dropped term, a wrong sign, index or
coefficient); each must make at least one test fail.  Two changes are
equivalent to the original and are expected to pass:
  * "lgr clamp z": a marker below the first cell centre gets cell 1 from the
    nearest-node search as well;
  * "solq3k dqb3n": the outflow Dirichlet value of the q3 k-lines only sets
    node n3 of the solve, which is not copied back (the n3m row does not
    couple to it: ap3sk(n3m) = 0), and q3(:,n3) gets dqb3n directly.

    python oracle/mutation_check.py [name-substring ...]
"""
import os
import shutil
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EQUIVALENT = {"lgr clamp z", "solq3k dqb3n"}

MUTATIONS = [
    ('prcalc sign',
     'PR(jc, kc) = PR(jc, kc) + DPH(jc, kc) -\n                         be *',
     'PR(jc, kc) = PR(jc, kc) + DPH(jc, kc) +\n                         be *'),
    ('pr_submean noop',
     's->pr[i] = s->pr[i] - pravg;',
     's->pr[i] = s->pr[i];'),
    ('boucdq 2.0',
     'dq2x2 = (s->qb2n[j] - Q2(j, n3m)) * 2.0 * s->udx3c[n3];',
     'dq2x2 = (s->qb2n[j] - Q2(j, n3m)) * s->udx3c[n3];'),
    ('cl 1/4',
     'lgr_h = lgr_h + 1.0 / 2.0 * philoc[i]',
     'lgr_h = lgr_h + 1.0 / 4.0 * philoc[i]'),
    ('solq2k outflow inj',
     'f[n3m] = f[n3m] + s->dqb2n[jc] * ugkkn;',
     'f[n3m] = f[n3m];'),
    ('helsolve fey 2dt',
     'RHS2(jc, kc) + FEY(jc, kc) * s->dt;',
     'RHS2(jc, kc) + FEY(jc, kc) * 2.0 * s->dt;'),
    ('hdnl2 h23 sign',
     'DPH(jc, kc) = DPH(jc, kc) - h23;\n        }\n    }',
     'DPH(jc, kc) = DPH(jc, kc) + h23;\n        }\n    }'),
    ('F13 zeroing',
     'FEY(j, s->n3m) = 0.0;\n            FEZ(j, s->n3m) = 0.0;',
     '(void)0;'),
    ('solq3j alpha',
     'static void solq3j(orc_state *s, double *dl, double *d, double *du, double *du2, int *ipiv, double *f,\n                   double *am, double *ap)\n{\n    int jc, kc, n2m = s->n2m, n3m = s->n3m;\n    double betadx = s->beta * s->al, b;',
     'static void solq3j(orc_state *s, double *dl, double *d, double *du, double *du2, int *ipiv, double *f,\n                   double *am, double *ap)\n{\n    int jc, kc, n2m = s->n2m, n3m = s->n3m;\n    double betadx = s->beta, b;'),
    ('invtr3 ro ru3',
     'RHS3(jc, kc) = (ga * QCAP(jc, kc) + ro * RU3(jc, kc) + alre * dcq3 - dpx33) * dt;\n            RU3(jc, kc) = QCAP(jc, kc);\n        }',
     'RHS3(jc, kc) = (ga * QCAP(jc, kc) + alre * dcq3 - dpx33) * dt;\n            RU3(jc, kc) = QCAP(jc, kc);\n        }'),
    ('invtr2 dp sign',
     'RHS2(jc, kc) = (ga * DPH(jc, kc) + ro * RU2(jc, kc) + alre * dcq2 - dpx22) * dt;',
     'RHS2(jc, kc) = (ga * DPH(jc, kc) + ro * RU2(jc, kc) + alre * dcq2 + dpx22) * dt;'),
    ('updvp udx3m',
     'usukm = al * dt * s->udx3c[kc];',
     'usukm = al * dt * s->udx3m[kc];'),
    ('hdnl3 h32 sign',
     'QCAP(jc, kc) = -(h32 + h33);',
     'QCAP(jc, kc) = -(-h32 + h33);'),
    ('hdnl3 h33 udx3m',
     '* s->udx3c[kc] * 0.25;\n            QCAP',
     '* s->udx3m[kc] * 0.25;\n            QCAP'),
    ('hdnl2 h22 coef',
     '* s->udx2c[jc] * 0.25;\n            DPH(jc, kc) = -h22;',
     '* s->udx2c[jc] * 0.5;\n            DPH(jc, kc) = -h22;'),
    ('hdnl2 k=n3m qb2n',
     'h23 = (s->qb2n[jc] * (s->qb3n[jc] + s->qb3n[jm]) -',
     'h23 = (Q2(jc, kc) * (s->qb3n[jc] + s->qb3n[jm]) -'),
    ('hdnl2 k=1 qb2s',
     's->qb2s[jc] * (s->qb3s[jc] + s->qb3s[jm])) * s->udx3m[kc] * 0.5;',
     's->qb2s[jc] * (s->qb3s[jc] + s->qb3s[jc])) * s->udx3m[kc] * 0.5;'),
    ('invtr2 k=1 codeb 2',
     '(Q2(jc, kc) - s->qb2s[jc]) * codeb * 2.0);',
     '(Q2(jc, kc) - s->qb2s[jc]) * codeb);'),
    ('invtr2 k=n3m 2.0',
     'd33q2 = ((s->qb2n[jc] - Q2(jc, kc)) * codea * 2.0',
     'd33q2 = ((s->qb2n[jc] - Q2(jc, kc)) * codea'),
    ('invtr3 q3lw',
     'dq32 = Q3(jp, kc) * appj + Q3(jc, kc) * accj + q3lw * ammj;',
     'dq32 = Q3(jp, kc) * appj + Q3(jc, kc) * accj;'),
    ('invtr3 q3uw',
     'dq32 = q3uw * appj + Q3(jc, kc) * accj',
     'dq32 = Q3(jc, kc) * accj'),
    ('invtr2 alre',
     'alre = al / s->ren;\n    for (kc = 2; kc <= n3m - 1; ++kc) {\n        km = s->kmv[kc]; kp = s->kpv[kc];',
     'alre = 1.0 / s->ren;\n    for (kc = 2; kc <= n3m - 1; ++kc) {\n        km = s->kmv[kc]; kp = s->kpv[kc];'),
    ('cfl 0.5',
     'fabs((Q3(j, k) + Q3(j, k + 1)) * 0.5 * udx3);',
     'fabs((Q3(j, k) + Q3(j, k + 1)) * udx3);'),
    ('cfl q3avg',
     'myq3avg = myq3avg + Q3(j, k) *',
     'myq3avg = myq3avg + Q2(j, k) *'),
    ('boucdq cor sign',
     's->dqb3n[j] = s->dqb3n[j] + cor;',
     's->dqb3n[j] = s->dqb3n[j] - cor;'),
    ('boucdq ro',
     's->dqb2n[j] = -dt * (ga * dq2x2 + ro * s->dq2x2o[j]) * cou;',
     's->dqb2n[j] = -dt * (ga * dq2x2) * cou;'),
    ('boucdq cou',
     's->dqb3n[j] = -dt * (ga * dq3x2 + ro * s->dq3x2o[j]) * cou;',
     's->dqb3n[j] = -dt * (ga * dq3x2 + ro * s->dq3x2o[j]);'),
    ('sd_ind n2m shift',
     'else if (indse[0] == n2m) jj = j - 1;',
     'else if (indse[0] == n2m) jj = j;'),
    ('sd_ind n3m shift',
     'else if (indse[1] == n3m) kk = k - 1;',
     'else if (indse[1] == n3m) kk = k;'),
    ('sd_ind k1 shift',
     'if (indse[1] == 1) kk = k + 1;',
     'if (indse[1] == 1) kk = k;'),
    ('sd_ind dsk',
     'dsk[1][m] = z[ind + 1] - z[ind];',
     'dsk[1][m] = z[ind] - z[ind - 1];'),
    ('gesolve no pivot',
     'if (fabs(a[j][k] / sc[j]) > pivot) {',
     'if (0) {'),
    ('wfun1 4rc3',
     'if (rcv <= 0.5) return (2.0 / 3.0) - 4.0 * rc2 + 4.0 * rc3;',
     'if (rcv <= 0.5) return (2.0 / 3.0) - 4.0 * rc2 + 3.0 * rc3;'),
    ('mls radius',
     'double sdds = radiusSD * sdd[k][i];',
     'double sdds = sdd[k][i];'),
    ('interp fcz sign',
     'fcz = (LMVEL(n, 2) - ucz) / s->dt;',
     'fcz = (ucz - LMVEL(n, 2)) / s->dt;'),
    ('interp fey cl',
     'FEY(j, k) = FEY(j, k) + CL(n, 1) * PHI(n, 1, m) * fcy;',
     'FEY(j, k) = FEY(j, k) + PHI(n, 1, m) * fcy;'),
    ('lgr clamp z',
     'if (z < s->zm[1] && z > 0.0) ind = 1;',
     'if (0) ind = 1;'),
    ('lgr minloc tie',
     'if (a < bv) { bv = a; best = i; }',
     'if (a <= bv) { bv = a; best = i; }'),
    ('divg usdtal',
     'double usdtal = 1.0 / (s->dt * s->al), dqcap;',
     'double usdtal = 1.0 / (s->dt), dqcap;'),
    ('phcalc wr sign',
     'double b = 1.0 / (s->acphj[j] + s->wr[k]);',
     'double b = 1.0 / (s->acphj[j] - s->wr[k]);'),
    ('updvp q2 alpha',
     'usujm = al * dt * s->udx2c[jc];',
     'usujm = dt * s->udx2c[jc];'),
    ('solq3k dqb3n',
     'f[n3] = s->dqb3n[jc];',
     'f[n3] = 0.0;'),
    ('solq3k q3 n3',
     'Q3(jc, n3) = Q3(jc, n3) + s->dqb3n[jc];',
     'Q3(jc, n3) = Q3(jc, n3);'),
    ('solq2k inflow ugkks',
     'ugkks = 1.0 / (s->g3rm[1] * s->g3rc[1]) * s->dx3q * 2.0 * betadx;',
     'ugkks = 1.0 / (s->g3rm[1] * s->g3rc[1]) * s->dx3q * betadx;'),
    ('solq2j beta',
     'b = 1.0 / (1.0 - betadx * s->ac2j[jc]);\n            ap[jc] = -betadx * s->ap2j[jc] * b;',
     'b = 1.0 / (1.0 - betadx * s->ac2j[jc]);\n            ap[jc] = -betadx * s->am2j[jc] * b;'),
    ('helsolve fez',
     'RHS3(jc, kc) + FEZ(jc, kc) * s->dt;',
     'RHS3(jc, kc);'),
    ('q2 wall zero',
     'Q2(n2, kc) = 0.0;\n        Q2(1, kc) = 0.0;',
     'Q2(1, kc) = 0.0;'),
    ('beta',
     's->beta = s->dt / s->ren * 0.5;',
     's->beta = s->dt / s->ren;'),
    ('boucqt qb2n',
     's->qb2n[j] = s->qb2n[j] + s->dqb2n[j];',
     's->qb2n[j] = s->qb2n[j];'),
    ('prcalc be',
     'double be = s->al * s->beta, ammk',
     'double be = s->beta, ammk'),
    ('submean area',
     'pravg = pravg / (s->p.rext2 * s->p.alx3);',
     'pravg = pravg / (s->p.rext2);'),
    ('dt max',
     'if (s->dt > s->p.dtmax) s->dt = s->p.dtmax;',
     '(void)0;'),
    ('AB rk ga',
     's->ga = 1.0 + 0.5 * s->dt / s->dt_o;',
     's->ga = 1.0 + s->dt / s->dt_o;'),
    ('AB ro',
     's->ro = -0.5 * s->dt / s->dt_o;',
     's->ro = -s->dt / s->dt_o;'),
    ('rk3 gam2',
     's->gam[2] = 5.0 / 12.0;',
     's->gam[2] = 5.0 / 11.0;'),
    ('rk3 rom3',
     's->rom[3] = -5.0 / 12.0;',
     's->rom[3] = -5.0 / 11.0;'),
    ('lgr clamp ztop',
     'else if (z > s->zm[n3m] && z < s->p.alx3) ind = n3m;',
     'else if (0) ind = n3m;'),
    ('lgr clamp ytop',
     'else if (y > s->rm[n2m] && y < s->p.rext2) ind = n2m;',
     'else if (0) ind = n2m;'),
    ('lgr outside',
     'if (ymax > s->p.rext2 || ymin < 0.0 || zmax > s->p.alx3 || zmin < 0.0) return ORC_ERR_MARKER_OUTSIDE;',
     '(void)0;'),
    ('pu check step',
     'for (j = 0; j < DIMSD; ++j) sumphi = sumphi + philoc[j];\n    if (fabs(1.0 - fabs(sumphi)) > 1.0e-10) return ORC_ERR_PARTITION_UNITY;',
     'for (j = 0; j < DIMSD; ++j) sumphi = sumphi + philoc[j];'),
    ('pu check loads',
     'for (i = 0; i < DIMSD; ++i) sumphi = sumphi + phid[0][i];\n    if (fabs(1.0 - fabs(sumphi)) > 1.0e-10) return ORC_ERR_PARTITION_UNITY;',
     'for (i = 0; i < DIMSD; ++i) sumphi = sumphi + phid[0][i];'),
    ('dt too small',
     'if (s->dt < 1.0e-8) {',
     'if (s->dt < 1.0e-30) {'),
    ('cfl lim',
     'if (s->cflm > s->p.cfllim) {',
     'if (0) {'),
    ('ns0 cflm',
     'if (ns == 0) {\n            s->cflm = s->cflm * s->dt;',
     'if (ns == 0) {\n            s->cflm = s->cflm;'),
    ('loads press 0.5',
     'press[n - 1] = pc[0] - 0.5 * (dpdnw + dpdne) * hpr;',
     'press[n - 1] = pc[0] - (dpdnw + dpdne) * hpr;'),
    ('loads dpdnw sign',
     'dpdnw = -1.0 * dpdnw;',
     'dpdnw = dpdnw;'),
    ('loads hpr',
     'const double hpr = ddmin * kprobe, visc',
     'const double hpr = ddmin, visc'),
    ('loads d3 0.5',
     '- 0.5 * d3ucdn3 * (hpr * hpr);',
     '- d3ucdn3 * (hpr * hpr);'),
    ('loads d2 sign',
     'const double ducdnm = ducdn - d2ucdn2 * hpr',
     'const double ducdnm = ducdn + d2ucdn2 * hpr'),
    ('loads tau visc',
     'tau[n - 1] = visc * ducdnm * lmtan[0];',
     'tau[n - 1] = ducdnm * lmtan[0];'),
    ('loads tan carry',
     'if (velt_md != 0.0) {',
     'if (1) {'),
    ('getpbup dz',
     'uk[2] = (f[GX(s, cy, cz)] - f[GX(s, cy, cz - 1)]) * s->dx3 / s->g3rc[cz];',
     'uk[2] = (f[GX(s, cy, cz)] - f[GX(s, cy, cz - 1)]) * s->dx3 / s->g3rc[cy];'),
    ('checkbcs norm',
     '*nbcv = sqrt(sv) / (double)nlm;',
     '*nbcv = sqrt(sv);'),
    ('checkbcs phipred',
     'ucz = ucz + PHIPRED(np, 2, n) * Q3(EULINS(np, 2, 1, n), EULINS(np, 2, 2, n));',
     'ucz = ucz + PHIPRED(np, 2, n) * Q3(EULINS(np, 2, 1, n), EULINS(np, 2, 2, n) + 1);'),
    ('diag ek 0.5',
     'ek = 0.5 * ek / (s->p.rext2 * s->p.alx3);',
     'ek = ek / (s->p.rext2 * s->p.alx3);'),
    ('diag vort sign',
     'vi = vi + (dwdy - dvdz) * (s->rc[j + 1] - s->rc[j]);',
     'vi = vi + (dwdy + dvdz) * (s->rc[j + 1] - s->rc[j]);'),
    ('diag vmin',
     'vmin3 = fmin(vmin3, Q3(j, k));',
     'vmin3 = fmin(vmin3, fabs(Q3(j, k)));'),
    ('diag div',
     'qmax = fmax(fabs(dq), qmax);',
     'qmax = fmax(dq, qmax);'),
    ('diag locz',
     'const double locz = 18.5;',
     'const double locz = 18.0;'),
]


def main(select):
    dst = tempfile.mkdtemp(prefix="fribo_mut_")
    shutil.copytree(ROOT, os.path.join(dst, "repo"),
                    ignore=shutil.ignore_patterns(".git", "_work", "gpurun_out", "__pycache__", "_build"))
    repo = os.path.join(dst, "repo")
    path = os.path.join(repo, "oracle", "fribo_oracle.c")
    orig = open(path).read()
    bad = 0
    try:
        for name, a, b in MUTATIONS:
            if select and not any(s in name for s in select):
                continue
            if orig.count(a) != 1:
                print(f"{name:24s} pattern not found once")
                bad += 1
                continue
            open(path, "w").write(orig.replace(a, b))
            subprocess.run(["make", "-s", "-B", "-C", os.path.dirname(path)], check=True,
                           stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
            r = subprocess.run([sys.executable, "-m", "pytest", "tests", "-q", "-x", "-m", "not gpu",
                                "-p", "no:cacheprovider"], cwd=repo, stdout=subprocess.PIPE,
                               stderr=subprocess.STDOUT, text=True)
            failed = [l.split(" ")[1] for l in r.stdout.splitlines() if l.startswith("FAILED")]
            caught = r.returncode != 0
            ok = caught != (name in EQUIVALENT)
            bad += not ok
            print(f"{name:24s} {'caught' if caught else 'passes'}{'' if ok else '  <-- NOT PINNED'}"
                  f"  {failed[0] if failed else ''}", flush=True)
    finally:
        shutil.rmtree(dst)
    return 1 if bad else 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
