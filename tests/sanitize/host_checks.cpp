// Host code of the library under AddressSanitizer and
// UndefinedBehaviorSanitizer (built and run by tests/test_host_sanitize.py):
//  * the exact decimal parser (psh_decimal.h, shared with the ingest
//    kernels) on edge-case and random tokens, checked against strtod;
//  * the Philox4x32-10 generator (philox.h) against the known-answer vectors
//    of Salmon et al. (2011);
//  * the output writers (psh_output.cpp) on a small survey and box result,
//    auto and cross, with and without the header.
// Any sanitizer report aborts the run (-fno-sanitize-recover=all).
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "philox.h"
#include "powspec_hip.h"
#include "psh_decimal.h"

// psh_save_res (in psh_output.cpp) reads a mesh's metadata through
// psh_mesh_info, which lives with the device code; these checks call
// psh_write_pk only, so the symbol is satisfied here and never reached.
extern "C" int psh_mesh_info(const psh_mesh *, int *, double *, double *, int *, int *) {
  return PSH_ERR_ARG;
}

static int failures = 0;
#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      failures++;                                                  \
    }                                                              \
  } while (0)

static long parsed = 0, declined = 0;

static void check_token(const std::string &t) {
  double got = 0.0;
  const bool ok = pshdec::parse_double(t.data(), t.data() + t.size(), &got);
  char *end = nullptr;
  const double want = std::strtod(t.c_str(), &end);
  if (!ok) {   // the parser may decline (the ingest then flags the line)
    declined++;
    return;
  }
  parsed++;
  if (end != t.c_str() + t.size()) {
    std::fprintf(stderr, "parsed a token strtod rejects: '%s'\n", t.c_str());
    failures++;
    return;
  }
  if (std::memcmp(&got, &want, sizeof got) != 0) {
    std::fprintf(stderr, "'%s': %.17g vs strtod %.17g\n", t.c_str(), got, want);
    failures++;
  }
}

static void decimal_checks() {
  const char *edge[] = {"0", "-0", "+0", "1", "-1", "0.1", "1e308", "1.7976931348623157e308",
                        "2.2250738585072014e-308", "4.9406564584124654e-324", "1e-320",
                        "123456789012345678901234567890", "0.000000000000000000000000001",
                        "9007199254740993", "9007199254740992.5", "1.5e-5", "3.14159e+2",
                        ".5", "5.", "1e0", "1E+05", "-2.5e-10", "", "-", "e5", "1e", "1.2.3",
                        "0x10", "nan", "inf", "1e400", "1e-400", "00000000000000000001.0"};
  for (const char *t : edge) check_token(t);
  std::mt19937_64 rng(12345);
  for (int n = 0; n < 200000; n++) {
    std::string t;
    if (rng() & 1) t += '-';
    const int nd = 1 + (int)(rng() % 40);
    const int dot = (int)(rng() % (nd + 1));
    for (int i = 0; i < nd; i++) {
      if (i == dot && (rng() & 1)) t += '.';
      t += (char)('0' + rng() % 10);
    }
    if (rng() % 3 == 0) {
      t += (rng() & 1) ? 'e' : 'E';
      const int e = (int)(rng() % 700) - 350;
      t += std::to_string(e);
    }
    check_token(t);
  }
  // shortest round-trip renderings of random doubles parse back to themselves
  for (int n = 0; n < 100000; n++) {
    uint64_t bits = rng();
    double v;
    std::memcpy(&v, &bits, sizeof v);
    if (!std::isfinite(v)) continue;
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.17g", v);
    check_token(buf);
  }
  // the comparisons above are not vacuous: nearly every number parses
  std::printf("decimal tokens: %ld parsed, %ld declined\n", parsed, declined);
  CHECK(parsed > 280000 && declined < 12000);
}

static void philox_checks() {
  // Salmon et al. (2011), Random123 known-answer vectors for Philox4x32-10
  const pshrng::W4 z = pshrng::philox4x32_10(0, 0, 0, 0, 0, 0);
  CHECK(z.v[0] == 0x6627e8d5u && z.v[1] == 0xe169c58du && z.v[2] == 0xbc57ac4cu &&
        z.v[3] == 0x9b00dbd8u);
  const pshrng::W4 o =
      pshrng::philox4x32_10(0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu,
                            0xffffffffu);
  CHECK(o.v[0] == 0x408f276du && o.v[1] == 0x41c83b0eu && o.v[2] == 0xa20bc7c6u &&
        o.v[3] == 0x6d5451fdu);
  const pshrng::W4 p = pshrng::philox4x32_10(0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u,
                                            0xa4093822u, 0x299f31d0u);
  CHECK(p.v[0] == 0xd16cfe09u && p.v[1] == 0x94fdccebu && p.v[2] == 0x5001e420u &&
        p.v[3] == 0x24126ea1u);
}

static void output_checks(const char *dir) {
  const int nb = 5, nl = 3;
  std::vector<double> k(nb), ke(nb + 1), km(nb), lc(nl * nb), p0(nl * nb), p1(nl * nb),
      xp(nl * nb);
  std::vector<uint64_t> cnt(nb);
  for (int i = 0; i <= nb; i++) ke[i] = 0.01 * i;
  for (int i = 0; i < nb; i++) {
    k[i] = 0.5 * (ke[i] + ke[i + 1]);
    km[i] = k[i] * 1.01;
    cnt[i] = 100 + 7 * i;
  }
  for (int i = 0; i < nl * nb; i++) {
    lc[i] = 1.0 + i;
    p0[i] = 1e4 / (i + 1);
    p1[i] = -2e3 / (i + 2);
    xp[i] = 5e2 * i;
  }
  psh_pk pk{};
  pk.nbin = nb;
  pk.nl = nl;
  pk.poles[0] = 0;
  pk.poles[1] = 2;
  pk.poles[2] = 4;
  pk.k = k.data();
  pk.kedge = ke.data();
  pk.km = km.data();
  pk.cnt = cnt.data();
  pk.lcnt = lc.data();
  pk.pl[0] = p0.data();
  pk.pl[1] = p1.data();
  pk.xpl = xp.data();
  psh_conf conf{};
  conf.ndata = 2;
  conf.gsize = 64;
  conf.assign = PSH_ASSIGN_TSC;
  conf.intlace = 1;
  conf.npole = nl;
  conf.poles[0] = 0;
  conf.poles[1] = 2;
  conf.poles[2] = 4;
  conf.kbin = 0.01;
  conf.isauto[0] = conf.isauto[1] = 1;
  conf.iscross = 1;
  size_t nd[2] = {1000, 2000}, nr[2] = {50000, 60000};
  double wd[2] = {1000, 2000}, wr[2] = {50000, 60000}, al[2] = {0.02, 0.033},
         sh[2] = {12.5, 7.25}, no[2] = {3.5, 4.25};
  psh_cata cat{};
  cat.num = 2;
  cat.ndata = nd;
  cat.nrand = nr;
  cat.wdata = wd;
  cat.wrand = wr;
  cat.alpha = al;
  cat.shot = sh;
  cat.norm = no;
  psh_mesh_meta meta{};
  meta.Ng = 64;
  meta.assign = PSH_ASSIGN_TSC;
  meta.intlace = 1;
  for (int a = 0; a < 3; a++) {
    meta.bsize[a] = 1000.0 + a;
    meta.bmin[a] = -500.0 + a;
  }
  const std::string fa0 = std::string(dir) + "/auto0.txt", fa1 = std::string(dir) + "/auto1.txt",
                    fx = std::string(dir) + "/cross.txt";
  const char *fauto[2] = {fa0.c_str(), fa1.c_str()};
  for (int issim = 0; issim <= 1; issim++)
    for (int header = 0; header <= 1; header++) {
      conf.issim = issim;
      CHECK(psh_write_pk(&conf, &cat, &meta, &pk, fauto, fx.c_str(), header) == PSH_OK);
      FILE *f = std::fopen(fa0.c_str(), "r");
      CHECK(f != nullptr);
      if (f) {
        int lines = 0;
        for (int ch; (ch = std::fgetc(f)) != EOF;) lines += ch == '\n';
        std::fclose(f);
        CHECK(lines >= nb);
      }
    }
  // argument errors come back as codes
  CHECK(psh_write_pk(nullptr, &cat, &meta, &pk, fauto, fx.c_str(), 1) == PSH_ERR_ARG);
  const char *bad[2] = {"", ""};
  CHECK(psh_write_pk(&conf, &cat, &meta, &pk, bad, fx.c_str(), 1) == PSH_ERR_ARG);
  const std::string nodir = std::string(dir) + "/no/such/dir/x.txt";
  const char *unwritable[2] = {nodir.c_str(), nodir.c_str()};
  CHECK(psh_write_pk(&conf, &cat, &meta, &pk, unwritable, fx.c_str(), 1) == PSH_ERR_FILE);
}

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <scratch dir>\n", argv[0]);
    return 2;
  }
  decimal_checks();
  philox_checks();
  output_checks(argv[1]);
  if (failures) {
    std::fprintf(stderr, "%d failures\n", failures);
    return 1;
  }
  std::printf("host checks passed\n");
  return 0;
}
