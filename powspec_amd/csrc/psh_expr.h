// psh_expr.h -- the ASCII ingest's column expressions: compiled on the host
// into a short postfix program, evaluated per line on the device.
//
// Grammar and meaning follow the reference's libast (lib/libast.c:79-201
// tokens and precedences, :797-858 tree insertion, :896-1138 tokenizer,
// :1467-1509 double evaluation, :1512-1855 boolean evaluation) as
// io/read_ascii.c:656-735 uses it:
//  * value expressions (positions, weights, n(z)) are double trees: every
//    column is converted to double and the operators are + - * / % (fmod)
//    ** (pow), unary -, abs( sqrt( ln( log( (log10) and parentheses;
//  * the selection is a boolean tree: comparisons < <= > >= == != on
//    numbers (a long operand meeting a double one is converted), ! on any
//    value, && || on booleans, isfinite( on doubles; numbers written with
//    digits only are long, the others double; %d columns are long.
// Binary operators associate to the left; unary - and ! bind tighter than
// every binary operator (-2**2 is 4), as in libast's insertion rule.
// Variables are $1 .. $9 and ${n}, 1-based among the formatter's value
// conversions.
//
// Evaluated on the device in double arithmetic; two longs divide with
// truncation, take C's remainder and libast's integer power, and the
// bitwise operators (<< >> & ^ | ~) take longs only.  String literals and
// comparisons mixing booleans with numbers are refused when the expression
// is compiled; long arithmetic agrees with libast while the values stay
// below 2^53 (an integer division by zero, which stops libast with SIGFPE,
// gives inf or nan here).
#pragma once

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

namespace pshexpr {

enum Op : int {
  OP_COL = 0,   // push value slot `arg`
  OP_IMM,       // push imm
  OP_NEG, OP_ABS, OP_SQRT, OP_LN, OP_LOG10, OP_ISFIN, OP_LNOT, OP_BNOT,
  OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_POW, OP_FMOD,
  OP_IDIV, OP_IMOD, OP_IPOW,   // two integers (libast's long arithmetic)
  OP_BAND, OP_BOR, OP_BXOR, OP_SHL, OP_SHR,
  OP_LT, OP_LE, OP_GT, OP_GE, OP_EQ, OP_NE, OP_AND, OP_OR
};

struct Ins {
  int op;
  int arg;
  double imm;
};

constexpr int kMaxIns = 64;     // instructions per expression
constexpr int kStack = 16;      // evaluation stack depth
constexpr int kMaxSlots = 16;   // distinct columns over all expressions of a file

// libast's power of two longs (lib/libast.c:1223-1290 and its caller,
// :1778-1795): exponents above 255 give 1, +-1 or 0; others are taken
// modulo 256 (a uint8_t) and from 63 on give the same overflow values;
// below, square-and-multiply in 64-bit two's complement.
__host__ __device__ inline long long ipow_long(long long base, long long e) {
  if (e <= 255) e = (unsigned char)e;
  if (e >= 63) return base == 1 ? 1 : (base == -1 ? 1 - 2 * (e & 1) : 0);
  unsigned long long r = 1, b = (unsigned long long)base;
  for (; e; e >>= 1, b *= b)
    if (e & 1) r *= b;
  return (long long)r;
}

// One program: n instructions from `ins`; the stack depth never exceeds
// kStack (checked when compiled).
__host__ __device__ inline double eval(const Ins *__restrict__ ins, int n, const double *slot) {
  double st[kStack];
  int sp = 0;
  for (int i = 0; i < n; i++) {
    const Ins I = ins[i];
    switch (I.op) {
      case OP_COL: st[sp++] = slot[I.arg]; break;
      case OP_IMM: st[sp++] = I.imm; break;
      case OP_NEG: st[sp - 1] = -st[sp - 1]; break;
      case OP_ABS: st[sp - 1] = fabs(st[sp - 1]); break;
      case OP_SQRT: st[sp - 1] = sqrt(st[sp - 1]); break;
      case OP_LN: st[sp - 1] = log(st[sp - 1]); break;
      case OP_LOG10: st[sp - 1] = log10(st[sp - 1]); break;
      case OP_ISFIN: st[sp - 1] = __builtin_isfinite(st[sp - 1]) ? 1.0 : 0.0; break;
      case OP_LNOT: st[sp - 1] = st[sp - 1] == 0.0 ? 1.0 : 0.0; break;
      case OP_BNOT: st[sp - 1] = (double)~(long long)st[sp - 1]; break;
      default: {
        const double b = st[--sp], a = st[sp - 1];
        double r;
        switch (I.op) {
          case OP_ADD: r = a + b; break;
          case OP_SUB: r = a - b; break;
          case OP_MUL: r = a * b; break;
          case OP_DIV: r = a / b; break;
          case OP_POW: r = pow(a, b); break;
          case OP_FMOD: r = fmod(a, b); break;
          // integers below 2^53: the rounded quotient truncates to C's
          case OP_IDIV: r = trunc(a / b); break;
          case OP_IMOD: r = fmod(a, b); break;   // exact, sign of a as C's %
          case OP_IPOW: r = (double)ipow_long((long long)a, (long long)b); break;
          // the bitwise operators of two longs; shift counts taken modulo 64
          // as the x86-64 shift instructions take them
          case OP_BAND: r = (double)((long long)a & (long long)b); break;
          case OP_BOR: r = (double)((long long)a | (long long)b); break;
          case OP_BXOR: r = (double)((long long)a ^ (long long)b); break;
          case OP_SHL:
            r = (double)(long long)((unsigned long long)(long long)a << ((long long)b & 63));
            break;
          case OP_SHR: r = (double)((long long)a >> ((long long)b & 63)); break;
          case OP_LT: r = a < b; break;
          case OP_LE: r = a <= b; break;
          case OP_GT: r = a > b; break;
          case OP_GE: r = a >= b; break;
          case OP_EQ: r = a == b; break;
          case OP_NE: r = a != b; break;
          case OP_AND: r = (a != 0.0) && (b != 0.0); break;
          default: r = (a != 0.0) || (b != 0.0); break;   // OP_OR
        }
        st[sp - 1] = r;
      }
    }
  }
  return st[0];
}

// ---- host compiler -----------------------------------------------------------
enum Type { T_DBL = 1, T_LONG = 2, T_BOOL = 4 };

struct Compiler {
  const char *src = nullptr, *p = nullptr;
  bool boolean = false;                  // selection (boolean tree)
  const std::vector<char> *integer = nullptr;   // per value column: %d-type
  std::vector<int> *slots = nullptr;     // value column (0-based) of each slot
  std::vector<Ins> out;
  int depth = 0, maxdepth = 0;
  std::string err;

  bool fail(const char *msg, const char *at) {
    if (err.empty()) {
      err = std::string(msg) + " at `" + std::string(at ? at : "") + "' in `" + src + "'";
    }
    return false;
  }
  void emit(int op, int arg = 0, double imm = 0.0) {
    out.push_back(Ins{op, arg, imm});
    if (op == OP_COL || op == OP_IMM) {
      if (++depth > maxdepth) maxdepth = depth;
    } else if (op >= OP_ADD) {
      depth--;
    }
  }
  void skip() {
    while (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\v' || *p == '\f') p++;
  }
  bool starts(const char *w) { return strncmp(p, w, strlen(w)) == 0; }

  // binary operator at p: its op code, precedence and length (0: none)
  int binop(int &op, int &len) {
    struct B { const char *s; int op, prec; };
    static const B tab[] = {
        {"**", OP_POW, 11}, {"*", OP_MUL, 10}, {"/", OP_DIV, 10}, {"%", OP_FMOD, 10},
        {"+", OP_ADD, 9},   {"-", OP_SUB, 9},  {"<<", OP_SHL, 8}, {">>", OP_SHR, 8},
        {"<=", OP_LE, 7},   {">=", OP_GE, 7},  {"<", OP_LT, 7},   {">", OP_GT, 7},
        {"==", OP_EQ, 6},   {"!=", OP_NE, 6},  {"&&", OP_AND, 2}, {"&", OP_BAND, 5},
        {"^", OP_BXOR, 4},  {"||", OP_OR, 1},  {"|", OP_BOR, 3}};
    for (const B &b : tab)
      if (starts(b.s)) {
        op = b.op;
        len = (int)strlen(b.s);
        return b.prec;
      }
    return 0;
  }

  // operand: unary operators, functions, parentheses, numbers, variables
  bool unary(int &t) {
    skip();
    const char *at = p;
    if (*p == '-' || *p == '!' || *p == '~') {
      const char u = *p;
      if (u == '!' && p[1] == '=') return fail("missing value", at);
      if (u != '-' && !boolean) return fail("invalid token for a value expression", at);
      p++;
      int a;
      if (!unary(a)) return false;
      if (u == '-') {
        if (a == T_BOOL) return fail("unexpected data type for operator", at);
        emit(OP_NEG);
        t = a;
      } else if (u == '~') {
        if (a != T_LONG) return fail("bitwise operator on a non-integer value", at);
        emit(OP_BNOT);
        t = T_LONG;
      } else {
        emit(OP_LNOT);
        t = T_BOOL;
      }
      return true;
    }
    struct F { const char *s; int op; };
    static const F fn[] = {{"abs(", OP_ABS}, {"sqrt(", OP_SQRT}, {"ln(", OP_LN},
                           {"log(", OP_LOG10}, {"isfinite(", OP_ISFIN}};
    for (const F &f : fn)
      if (starts(f.s)) {
        if (f.op == OP_ISFIN && !boolean)
          return fail("invalid token for a value expression", at);
        p += strlen(f.s);
        int a;
        if (!expr(1, a)) return false;
        skip();
        if (*p != ')') return fail("unclosed parenthesis", p);
        p++;
        if (a == T_BOOL) return fail("unexpected data type for operator", at);
        if (f.op == OP_ISFIN && a != T_DBL)
          return fail("isfinite of an integer value", at);
        emit(f.op);
        t = f.op == OP_ABS ? a : (f.op == OP_ISFIN ? T_BOOL : T_DBL);
        return true;
      }
    if (*p == '(') {
      p++;
      skip();
      if (*p == ')') return fail("empty parenthesis", p);
      if (!expr(1, t)) return false;
      skip();
      if (*p != ')') return fail("unclosed parenthesis", p);
      p++;
      return true;
    }
    if ((*p >= '0' && *p <= '9') || *p == '.') {
      char *end = nullptr;
      const double v = strtod(p, &end);
      if (end == p) return fail("unrecognised number", at);
      if (v == HUGE_VAL) return fail("overflow detected for double type", at);
      t = T_DBL;
      if (boolean) {   // digits only: a long literal
        bool digits = true;
        for (const char *q = p; q < end; q++) digits = digits && (*q >= '0' && *q <= '9');
        if (digits) t = T_LONG;
      }
      p = end;
      emit(OP_IMM, 0, v);
      return true;
    }
    if (*p == '$') {
      long idx = 0;
      if (p[1] >= '1' && p[1] <= '9') {
        idx = p[1] - '0';
        p += 2;
      } else if (p[1] == '{') {
        const char *q = p + 2;
        while (*q >= '0' && *q <= '9' && idx < 100000) idx = idx * 10 + (*q++ - '0');
        if (idx <= 0 || *q != '}') return fail("unrecognised variable", at);
        p = q + 1;
      } else {
        return fail("unrecognised variable", at);
      }
      if (idx > (long)integer->size())
        return fail("the column is beyond the formatter's value columns", at);
      const int col = (int)idx - 1;
      int s = -1;
      for (size_t i = 0; i < slots->size(); i++)
        if ((*slots)[i] == col) s = (int)i;
      if (s < 0) {
        if ((int)slots->size() >= kMaxSlots)
          return fail("more than 16 distinct columns in the expressions", at);
        s = (int)slots->size();
        slots->push_back(col);
      }
      emit(OP_COL, s);
      t = boolean && (*integer)[col] ? T_LONG : T_DBL;
      return true;
    }
    if (*p == '"' || *p == '\'') return fail("string literals are not supported", at);
    if (!*p || *p == ')') return fail("incomplete expression", at);
    return fail("unrecognised token", at);
  }

  // precedence climbing, every binary operator left-associative
  bool expr(int minprec, int &t) {
    if (!unary(t)) return false;
    for (;;) {
      skip();
      int op, len;
      const char *at = p;
      const int prec = binop(op, len);
      if (prec == 0 || prec < minprec) return true;
      const bool cmp = op >= OP_LT && op <= OP_NE, logic = op == OP_AND || op == OP_OR;
      const bool bits = op >= OP_BAND && op <= OP_SHR;
      if (!boolean && (cmp || logic || bits))
        return fail("invalid token for a value expression", at);
      p += len;
      int r;
      if (!expr(prec + 1, r)) return false;
      if (logic) {
        if (t != T_BOOL || r != T_BOOL) return fail("unexpected data type for operator", at);
        t = T_BOOL;
      } else if (bits) {
        if (t != T_LONG || r != T_LONG)
          return fail("bitwise operator on a non-integer value", at);
      } else if (op == OP_EQ || op == OP_NE) {
        if ((t == T_BOOL) != (r == T_BOOL))
          return fail("comparison of a boolean with a number", at);
        t = T_BOOL;
      } else {
        if (t == T_BOOL || r == T_BOOL) return fail("unexpected data type for operator", at);
        if (cmp) {
          t = T_BOOL;
        } else {
          if (t == T_LONG && r == T_LONG) {
            if (op == OP_DIV) op = OP_IDIV;
            else if (op == OP_FMOD) op = OP_IMOD;
            else if (op == OP_POW) op = OP_IPOW;
          }
          t = (t == T_LONG && r == T_LONG) ? T_LONG : T_DBL;
        }
      }
      emit(op);
    }
  }

  // Whole expression; false with err set when it does not compile.
  bool compile(const char *s) {
    src = p = s;
    skip();
    if (!*p) return fail("empty expression", p);
    int t;
    if (!expr(1, t)) return false;
    skip();
    if (*p == ')') return fail("unbalanced parenthesis", p);
    if (*p) return fail("missing operator", p);
    if (boolean && t != T_BOOL) return fail("the selection is not a boolean expression", s);
    if ((int)out.size() > kMaxIns) return fail("expression longer than 64 operations", s);
    if (maxdepth > kStack) return fail("expression nested deeper than 16 values", s);
    return true;
  }
};

}  // namespace pshexpr
