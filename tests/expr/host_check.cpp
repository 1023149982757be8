// Host build of the ASCII ingest's expression compiler and evaluator
// (powspec_amd/csrc/psh_expr.h), for the fuzz test against the reference's
// libast: test infrastructure only.
//
// stdin: ncase, then ncase lines "D <expr>" / "B <expr>", then nrow and
// nrow rows of the six value columns of formatter "%lf %d %lf %lf %d %lf".
// stdout per case: "REFUSED <message>" or "OK" followed by one value per
// row (%a for values, 0/1 for selections).
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "../../powspec_amd/csrc/psh_expr.h"

using namespace pshexpr;

int main() {
  std::string line;
  if (!std::getline(std::cin, line)) return 1;
  const int ncase = atoi(line.c_str());
  std::vector<char> kind;
  std::vector<std::string> src;
  for (int i = 0; i < ncase; i++) {
    if (!std::getline(std::cin, line) || line.size() < 2) return 1;
    kind.push_back(line[0]);
    src.push_back(line.substr(2));
  }
  if (!std::getline(std::cin, line)) return 1;
  const int nrow = atoi(line.c_str());
  std::vector<double> rows((size_t)nrow * 6);
  for (double &v : rows)
    if (!(std::cin >> v)) return 1;
  const std::vector<char> integer = {0, 1, 0, 0, 1, 0};
  for (int i = 0; i < ncase; i++) {
    std::vector<int> slots;
    Compiler cp;
    cp.boolean = kind[i] == 'B';
    cp.integer = &integer;
    cp.slots = &slots;
    if (!cp.compile(src[i].c_str())) {
      printf("REFUSED %s\n", cp.err.c_str());
      continue;
    }
    printf("OK");
    for (int r = 0; r < nrow; r++) {
      double s[kMaxSlots];
      for (size_t k = 0; k < slots.size(); k++) s[k] = rows[(size_t)r * 6 + slots[k]];
      const double v = eval(cp.out.data(), (int)cp.out.size(), s);
      if (cp.boolean) printf(" %d", v != 0.0 ? 1 : 0);
      else printf(" %a", v);
    }
    printf("\n");
  }
  return 0;
}
