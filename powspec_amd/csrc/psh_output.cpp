// psh_output.cpp -- the output files of the reference's save_res
// (src/save_res.c:35-230), written from the psh_pk arrays.  Host code only.
#include <stdio.h>
#include <string.h>

#include "powspec_hip.h"

namespace {

const char *const kAssignNames[] = {"NGP", "CIC", "TSC", "PCS"};   // src/genr_mesh.c:27-32
const char kComment = '#';                                       // src/define.h:87
#define PSH_OFMT "%.10lg"                                       // src/define.h:107

struct File {
  FILE *fp = nullptr;
  bool ok = true;
  explicit File(const char *name) : fp(fopen(name, "w")) { ok = fp != nullptr; }
  ~File() {
    if (fp && fclose(fp)) ok = false;
  }
  template <class... A>
  void put(const char *fmt, A... a) {
    if constexpr (sizeof...(A) == 0) {
      if (ok && fputs(fmt, fp) < 0) ok = false;
    } else {
      if (ok && fprintf(fp, fmt, a...) < 0) ok = false;
    }
  }
};

void mesh_lines(File &f, const psh_mesh_meta *m) {
  f.put("%c Box size: [%lg, %lg, %lg]\n%c Box boundaries: [[%lg,%lg], [%lg,%lg], [%lg,%lg]]\n",
        kComment, m->bsize[0], m->bsize[1], m->bsize[2], kComment, m->bmin[0],
        m->bmin[0] + m->bsize[0], m->bmin[1], m->bmin[1] + m->bsize[1], m->bmin[2],
        m->bmin[2] + m->bsize[2]);
  f.put("%c Grid size: %d , assignment scheme: %s , grid interlacing: %s\n", kComment, m->Ng,
        kAssignNames[m->assign & 3], m->intlace ? "enabled" : "disabled");
}

void columns(File &f, const psh_pk *pk, const double *pl) {
  f.put("%c kcen(1) kmin(2) kmax(3) kavg(4) nmod(5)", kComment);
  for (int l = 0; l < pk->nl; l++) f.put(" P_%d(%d)", pk->poles[l], l + 6);
  f.put("\n");
  for (int i = 0; i < pk->nbin; i++) {
    f.put(PSH_OFMT " " PSH_OFMT " " PSH_OFMT " " PSH_OFMT " %zu", pk->k[i], pk->kedge[i],
          pk->kedge[i + 1], pk->km[i], (size_t)pk->cnt[i]);
    for (int l = 0; l < pk->nl; l++) f.put(" " PSH_OFMT, pl[(size_t)l * pk->nbin + i]);
    f.put("\n");
  }
}

}  // namespace

extern "C" int psh_write_pk(const psh_conf *conf, const psh_cata *cat, const psh_mesh_meta *m,
                            const psh_pk *pk, const char *const fauto[2], const char *fcross,
                            int header) {
  if (!conf || !cat || !m || !pk) return PSH_ERR_ARG;
  for (int n = 0; n < cat->num && n < 2; n++) {
    if (!conf->isauto[n]) continue;
    if (!fauto || !fauto[n] || !*fauto[n] || !pk->pl[n]) return PSH_ERR_ARG;
    File f(fauto[n]);
    if (!f.ok) return PSH_ERR_FILE;
    if (header) {
      f.put("%c Data catalog: %zu objects, total weight: " PSH_OFMT "\n", kComment,
            cat->ndata[n], cat->wdata[n]);
      if (!conf->issim)
        f.put("%c Random catalog: %zu objects, total weight: " PSH_OFMT "\n", kComment,
              cat->nrand[n], cat->wrand[n]);
      mesh_lines(f, m);
      // surveys report shot / norm (src/save_res.c:87-90)
      f.put("%c Shot noise: " PSH_OFMT " , normalisation: " PSH_OFMT " \n", kComment,
            conf->issim ? cat->shot[n] : cat->shot[n] / cat->norm[n], cat->norm[n]);
    }
    columns(f, pk, pk->pl[n]);
    if (!f.ok) return PSH_ERR_FILE;
  }
  if (!conf->iscross) return PSH_OK;
  if (!fcross || !*fcross || !pk->xpl || cat->num < 2) return PSH_ERR_ARG;
  File f(fcross);
  if (!f.ok) return PSH_ERR_FILE;
  if (header) {
    f.put("%c Data catalog 1: %zu objects, total weight: " PSH_OFMT
          "\n%c Data catalog 2: %zu objects, total weight: " PSH_OFMT "\n",
          kComment, cat->ndata[0], cat->wdata[0], kComment, cat->ndata[1], cat->wdata[1]);
    if (!conf->issim)
      f.put("%c Random catalog 1: %zu objects, total weight: " PSH_OFMT
            "\n%c Random catalog 2: %zu objects, total weight: " PSH_OFMT "\n",
            kComment, cat->nrand[0], cat->wrand[0], kComment, cat->nrand[1], cat->wrand[1]);
    mesh_lines(f, m);
    f.put("%c Normalisation factor 1: " PSH_OFMT " , Normalisation factor 2: " PSH_OFMT "\n",
          kComment, cat->norm[0], cat->norm[1]);
  }
  columns(f, pk, pk->xpl);
  return f.ok ? PSH_OK : PSH_ERR_FILE;
}

extern "C" int psh_save_res(const psh_conf *conf, const psh_cata *cat, const psh_mesh *mesh,
                            const psh_pk *pk, const char *const fauto[2], const char *fcross,
                            int header) {
  psh_mesh_meta m;
  memset(&m, 0, sizeof m);
  if (!mesh || psh_mesh_info(mesh, &m.Ng, m.bsize, m.bmin, &m.assign, &m.intlace))
    return PSH_ERR_ARG;
  return psh_write_pk(conf, cat, &m, pk, fauto, fcross, header);
}
