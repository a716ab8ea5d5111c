// Expression programs: the initial and boundary conditions of the multigrid
// front end written as evalexpr expressions (multigrid/src/eval/evalexpr.rs;
// the `evalexpr` crate, version 7.2.0, evaluates them there), compiled on the
// host (sbp_expr.cpp) into a small stack program that runs per point on the
// device (far-field edge data at every stage, fields for initial conditions
// and errors) and on the host (the `globals` statements, once).
//
// Values carry evalexpr's dynamic type: integers (kept as exact doubles),
// floats and booleans. Integer arithmetic stays integer (5/2 = 2, 7 % 3 = 1),
// mixed arithmetic is float, `^` is always float, comparisons give booleans.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sbp {
namespace expr {

enum Tag : int32_t { TAG_INT = 0, TAG_FLOAT = 1, TAG_BOOL = 2, TAG_EMPTY = 3 };

enum Op : int32_t {
  OP_CONST = 0,  // push consts[arg]
  OP_INPUT,      // push input arg (Input) as a float
  OP_LOAD,       // push slot arg
  OP_STORE,      // pop into slot arg, push empty (an assignment's value)
  OP_POP,        // discard the top (`;` between statements)
  OP_EMPTY,      // push empty (an empty statement after `;`)
  OP_ADD,
  OP_SUB,
  OP_MUL,
  OP_DIV,
  OP_MOD,
  OP_POW,
  OP_LT,
  OP_GT,
  OP_LE,
  OP_GE,
  OP_EQ,
  OP_NE,
  OP_AND,
  OP_OR,
  OP_NEG,
  OP_NOT,
  OP_CALL,  // arg = function | nargs << 8
};

// Inputs of an expression (ContextWrapper::get_value, evalexpr.rs:141-154).
enum Input : int32_t { IN_T = 0, IN_X, IN_Y, IN_RHO, IN_RHOU, IN_RHOV, IN_U, IN_V, IN_COUNT };

enum Fn : int32_t {
  FN_IF = 0,  // the configuration's own `if` and `case` (evalexpr.rs:307-348)
  FN_CASE,
  FN_MIN,
  FN_MAX,
  FN_FLOOR,
  FN_ROUND,
  FN_CEIL,
  FN_ABS,
  FN_LN,
  FN_LOG,
  FN_LOG2,
  FN_LOG10,
  FN_EXP,
  FN_EXP2,
  FN_POW,
  FN_SQRT,
  FN_CBRT,
  FN_HYPOT,
  FN_SIN,
  FN_COS,
  FN_TAN,
  FN_ASIN,
  FN_ACOS,
  FN_ATAN,
  FN_ATAN2,
  FN_SINH,
  FN_COSH,
  FN_TANH,
  FN_ASINH,
  FN_ACOSH,
  FN_ATANH,
  FN_IS_NAN,
  FN_IS_FINITE,
  FN_IS_INFINITE,
  FN_IS_NORMAL,
  FN_COUNT
};

// The two forms of multigrid/src/input.rs:165-190: conservative variables
// (rho, rhou, rhov, e) or primitive ones (rho, u, v, p).
enum Form : int32_t { FORM_CONSERVATION = 0, FORM_PRESSURE = 1 };

constexpr int kMaxStack = 32;
constexpr int kMaxSlots = 16;
constexpr int kMaxArgs = 15;

struct Instr {
  int32_t op, arg;
};

struct Value {
  double v;
  int32_t tag, pad_;
};

// One compiled form (four expressions) as one contiguous blob: this header,
// then `ncode` instructions, then `nconst` constants. The four expressions
// start at code offsets off[i] and run len[i] instructions.
struct Program {
  int32_t form, ncode, nconst, pad_;
  int32_t off[4], len[4];
  double gamma;
};

__host__ __device__ inline const Instr* code_of(const Program* P) { return reinterpret_cast<const Instr*>(P + 1); }
__host__ __device__ inline const Value* consts_of(const Program* P) {
  return reinterpret_cast<const Value*>(code_of(P) + P->ncode);
}
inline size_t program_bytes(int ncode, int nconst) {
  size_t b = sizeof(Program) + sizeof(Instr) * (size_t)ncode;
  b = (b + 15) / 16 * 16;
  return b + sizeof(Value) * (size_t)nconst;
}

__host__ __device__ inline bool is_num(const Value& a) { return a.tag == TAG_INT || a.tag == TAG_FLOAT; }
__host__ __device__ inline Value mk(double v, int32_t tag) {
  Value r;
  r.v = v;
  r.tag = tag;
  r.pad_ = 0;
  return r;
}
// integers must stay exact doubles (evalexpr's checked i64 arithmetic fails
// on overflow; beyond 2^53 this treats it as a failure too)
__host__ __device__ inline bool int_ok(double v) { return v >= -9007199254740992.0 && v <= 9007199254740992.0; }

// Runs the expression at code[0..len) with the given inputs; false on an
// evaluation error (a type error, an integer division by zero, a missing
// value), which the reference reports by panicking.
__host__ __device__ inline bool run(const Instr* code, int len, const Value* consts, const double* in, Value& out,
                                    Value* slots_out = nullptr) {
  Value st[kMaxStack];
  Value slot[kMaxSlots];
  int sp = 0;
  for (int s = 0; s < kMaxSlots; ++s) slot[s] = mk(0.0, TAG_EMPTY);
  for (int pc = 0; pc < len; ++pc) {
    const Instr I = code[pc];
    switch (I.op) {
      case OP_CONST: st[sp++] = consts[I.arg]; break;
      case OP_INPUT: st[sp++] = mk(in[I.arg], TAG_FLOAT); break;
      case OP_LOAD:
        if (slot[I.arg].tag == TAG_EMPTY) return false;
        st[sp++] = slot[I.arg];
        break;
      case OP_STORE: slot[I.arg] = st[--sp]; st[sp++] = mk(0.0, TAG_EMPTY); break;
      case OP_POP: --sp; break;
      case OP_EMPTY: st[sp++] = mk(0.0, TAG_EMPTY); break;
      case OP_NEG: {
        Value& a = st[sp - 1];
        if (!is_num(a)) return false;
        a.v = -a.v;
        break;
      }
      case OP_NOT: {
        Value& a = st[sp - 1];
        if (a.tag != TAG_BOOL) return false;
        a.v = a.v != 0.0 ? 0.0 : 1.0;
        break;
      }
      case OP_AND:
      case OP_OR: {
        const Value b = st[--sp];
        Value& a = st[sp - 1];
        if (a.tag != TAG_BOOL || b.tag != TAG_BOOL) return false;
        a.v = I.op == OP_AND ? ((a.v != 0.0 && b.v != 0.0) ? 1.0 : 0.0) : ((a.v != 0.0 || b.v != 0.0) ? 1.0 : 0.0);
        break;
      }
      case OP_EQ:
      case OP_NE: {
        const Value b = st[--sp];
        Value& a = st[sp - 1];
        bool eq;
        if (is_num(a) && is_num(b))
          eq = a.tag == b.tag ? a.v == b.v : false;  // Value::eq: Int(1) != Float(1.0)
        else
          eq = a.tag == b.tag && a.v == b.v;
        a = mk((I.op == OP_EQ) == eq ? 1.0 : 0.0, TAG_BOOL);
        break;
      }
      case OP_LT:
      case OP_GT:
      case OP_LE:
      case OP_GE: {
        const Value b = st[--sp];
        Value& a = st[sp - 1];
        if (!is_num(a) || !is_num(b)) return false;
        bool r = I.op == OP_LT ? a.v < b.v : I.op == OP_GT ? a.v > b.v : I.op == OP_LE ? a.v <= b.v : a.v >= b.v;
        a = mk(r ? 1.0 : 0.0, TAG_BOOL);
        break;
      }
      case OP_ADD:
      case OP_SUB:
      case OP_MUL:
      case OP_DIV:
      case OP_MOD:
      case OP_POW: {
        const Value b = st[--sp];
        Value& a = st[sp - 1];
        if (!is_num(a) || !is_num(b)) return false;
        if (I.op == OP_POW) {
          a = mk(pow(a.v, b.v), TAG_FLOAT);
        } else if (a.tag == TAG_INT && b.tag == TAG_INT) {
          double r;
          switch (I.op) {
            case OP_ADD: r = a.v + b.v; break;
            case OP_SUB: r = a.v - b.v; break;
            case OP_MUL: r = a.v * b.v; break;
            case OP_DIV:
              if (b.v == 0.0) return false;
              r = trunc(a.v / b.v);
              break;
            default:
              if (b.v == 0.0) return false;
              r = fmod(a.v, b.v);  // the sign of the dividend, as Rust's %
          }
          if (!int_ok(r)) return false;
          a.v = r;
        } else {
          switch (I.op) {
            case OP_ADD: a.v = a.v + b.v; break;
            case OP_SUB: a.v = a.v - b.v; break;
            case OP_MUL: a.v = a.v * b.v; break;
            case OP_DIV: a.v = a.v / b.v; break;
            default: a.v = fmod(a.v, b.v);
          }
          a.tag = TAG_FLOAT;
        }
        break;
      }
      case OP_CALL: {
        const int fn = I.arg & 0xff, na = I.arg >> 8;
        Value* a = st + sp - na;
        Value r;
        switch (fn) {
          case FN_IF:
            if (a[0].tag != TAG_BOOL) return false;
            r = a[0].v != 0.0 ? a[1] : a[2];
            break;
          case FN_CASE: {
            r = a[na - 1];
            for (int k = 0; k + 1 < na; k += 2) {
              if (a[k].tag != TAG_BOOL) return false;
              if (a[k].v != 0.0) {
                r = a[k + 1];
                break;
              }
            }
            break;
          }
          case FN_MIN:
          case FN_MAX: {
            bool all_int = true;
            for (int k = 0; k < na; ++k) {
              if (!is_num(a[k])) return false;
              all_int = all_int && a[k].tag == TAG_INT;
            }
            double m = a[0].v;
            for (int k = 1; k < na; ++k) m = fn == FN_MIN ? fmin(m, a[k].v) : fmax(m, a[k].v);
            r = mk(m, all_int ? TAG_INT : TAG_FLOAT);
            break;
          }
          case FN_FLOOR:
          case FN_ROUND:
          case FN_CEIL:
          case FN_ABS:
            if (!is_num(a[0])) return false;
            if (a[0].tag == TAG_INT) {
              r = fn == FN_ABS ? mk(fabs(a[0].v), TAG_INT) : a[0];
            } else {
              // f64::round rounds half away from zero, as round()
              const double v = a[0].v;
              r = mk(fn == FN_FLOOR ? floor(v) : fn == FN_ROUND ? round(v) : fn == FN_CEIL ? ceil(v) : fabs(v), TAG_FLOAT);
            }
            break;
          case FN_IS_NAN:
          case FN_IS_FINITE:
          case FN_IS_INFINITE:
          case FN_IS_NORMAL: {
            if (!is_num(a[0])) return false;
            const double v = a[0].v;
            bool b = fn == FN_IS_NAN ? __builtin_isnan(v) : fn == FN_IS_FINITE ? __builtin_isfinite(v) : fn == FN_IS_INFINITE ? __builtin_isinf(v)
                                                                                                          : __builtin_isnormal(v);
            r = mk(b ? 1.0 : 0.0, TAG_BOOL);
            break;
          }
          default: {
            for (int k = 0; k < na; ++k)
              if (!is_num(a[k])) return false;
            const double u = a[0].v, w = na > 1 ? a[1].v : 0.0;
            double v;
            switch (fn) {
              case FN_LN: v = log(u); break;
              case FN_LOG: v = log(u) / log(w); break;  // f64::log(self, base)
              case FN_LOG2: v = log2(u); break;
              case FN_LOG10: v = log10(u); break;
              case FN_EXP: v = exp(u); break;
              case FN_EXP2: v = exp2(u); break;
              case FN_POW: v = pow(u, w); break;
              case FN_SQRT: v = sqrt(u); break;
              case FN_CBRT: v = cbrt(u); break;
              case FN_HYPOT: v = hypot(u, w); break;
              case FN_SIN: v = sin(u); break;
              case FN_COS: v = cos(u); break;
              case FN_TAN: v = tan(u); break;
              case FN_ASIN: v = asin(u); break;
              case FN_ACOS: v = acos(u); break;
              case FN_ATAN: v = atan(u); break;
              case FN_ATAN2: v = atan2(u, w); break;
              case FN_SINH: v = sinh(u); break;
              case FN_COSH: v = cosh(u); break;
              case FN_TANH: v = tanh(u); break;
              case FN_ASINH: v = asinh(u); break;
              case FN_ACOSH: v = acosh(u); break;
              default: v = atanh(u); break;
            }
            r = mk(v, TAG_FLOAT);
          }
        }
        sp -= na;
        st[sp++] = r;
        break;
      }
      default: return false;
    }
  }
  if (sp != 1) return false;
  out = st[0];
  if (slots_out)
    for (int s = 0; s < kMaxSlots; ++s) slots_out[s] = slot[s];
  return true;
}

// The four conservative variables of one point at time t (the evaluators of
// multigrid/src/eval/evalexpr.rs:168-208 and, for the pressure form,
// euler/src/eval.rs's EvaluatorPressureWrapper: rho, u, v, p, then
// e = p/(γ-1) + rho(u² + v²)/2, rhou = rho·u, rhov = rho·v). Each expression
// value must be a number (eval_number); false on any evaluation error.
__host__ __device__ inline bool eval_point(const Program* P, double t, double x, double y, double q[4]) {
#pragma clang fp contract(off)
  const Instr* code = code_of(P);
  const Value* cs = consts_of(P);
  double in[IN_COUNT] = {t, x, y, 0.0, 0.0, 0.0, 0.0, 0.0};
  double v[4];
  for (int i = 0; i < 4; ++i) {
    Value r;
    if (!run(code + P->off[i], P->len[i], cs, in, r) || !is_num(r)) return false;
    v[i] = r.v;
    if (P->form == FORM_CONSERVATION) {
      // rhou and rhov see rho; e sees rho, rhou and rhov
      if (i == 0) in[IN_RHO] = v[0];
      if (i == 1) in[IN_RHOU] = v[1];
      if (i == 2) in[IN_RHOV] = v[2];
    } else {
      // u and v see rho; p sees rho, u and v
      if (i == 0) in[IN_RHO] = v[0];
      if (i == 1) in[IN_U] = v[1];
      if (i == 2) in[IN_V] = v[2];
    }
  }
  if (P->form == FORM_CONSERVATION) {
    for (int c = 0; c < 4; ++c) q[c] = v[c];
  } else {
    const double rho = v[0], u = v[1], w = v[2], p = v[3];
    q[0] = rho;
    q[3] = p / (P->gamma - 1.0) + rho * (u * u + w * w) / 2.0;
    q[1] = u * rho;
    q[2] = w * rho;
  }
  return true;
}

}  // namespace expr
}  // namespace sbp
