// Host compiler of expression programs (sbp_expr.hpp): the evalexpr syntax the
// multigrid configurations use for initial and boundary conditions
// (multigrid/src/eval/evalexpr.rs builds operator trees with the `evalexpr`
// crate, version 7.2.0, and evaluates them per point), restated as a
// tokenizer, a precedence-climbing parser and an emitter of stack programs.
//
// Operators by precedence (evalexpr 7's table), all left-associative except
// assignment: `;` 0, `,` 40 (tuples, as function arguments), `=` and the
// compound assignments 50, `||` 70, `&&` 75, comparisons 80, `+ -` 95,
// `* / %` 100, unary `- !` 110, `^` 120, function application 190.
// Names resolve as ContextWrapper::get_value does (evalexpr.rs:141-154): the
// point's inputs t, x, y (and rho, rhou, rhov, u, v where the form provides
// them) first, then variables the expression assigned, then the globals
// (GAMMA, PI, E and whatever the `globals` statements assigned, evaluated
// once at compile time, evalexpr.rs:18-73).
// The reference keeps an expression's assigned variables from one point to
// the next; an expression that reads such a variable before assigning it
// would see the previous point's value. Those expressions are refused here,
// and so are unknown names, tuples outside function arguments, unknown
// functions and wrong argument counts, all of which the reference panics on.
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "../../include/sbpgpu.h"
#include "sbp_expr.hpp"

using namespace sbp::expr;

namespace {

struct Tok {
  enum Kind { NUM, IDENT, OP, END } kind;
  std::string s;
  size_t pos;
};

struct Node {
  enum Kind { NUM, BOOL, VAR, BIN, NEG, NOT, ASSIGN, CALL, TUPLE, CHAIN, EMPTY } kind;
  Value val{};
  std::string name;  // VAR, ASSIGN target, CALL function, BIN / ASSIGN operator
  std::vector<std::unique_ptr<Node>> kids;
};
using NodeP = std::unique_ptr<Node>;

struct Error {
  std::string msg;
};

bool is_ident_start(char c) { return std::isalpha((unsigned char)c) || c == '_'; }
bool is_ident_part(char c) { return std::isalnum((unsigned char)c) || c == '_' || c == ':'; }

std::vector<Tok> tokenize(const std::string& src) {
  static const char* ops3[] = {"&&=", "||="};
  static const char* ops2[] = {"==", "!=", "<=", ">=", "&&", "||", "+=", "-=", "*=", "/=", "%=", "^="};
  std::vector<Tok> out;
  size_t i = 0;
  while (i < src.size()) {
    const char c = src[i];
    if (std::isspace((unsigned char)c)) {
      ++i;
      continue;
    }
    if (std::isdigit((unsigned char)c) || (c == '.' && i + 1 < src.size() && std::isdigit((unsigned char)src[i + 1]))) {
      size_t j = i;
      while (j < src.size() && (std::isdigit((unsigned char)src[j]) || src[j] == '.')) ++j;
      if (j < src.size() && (src[j] == 'e' || src[j] == 'E')) {
        size_t k = j + 1;
        if (k < src.size() && (src[k] == '+' || src[k] == '-')) ++k;
        if (k < src.size() && std::isdigit((unsigned char)src[k])) {
          while (k < src.size() && std::isdigit((unsigned char)src[k])) ++k;
          j = k;
        }
      }
      out.push_back({Tok::NUM, src.substr(i, j - i), i});
      i = j;
      continue;
    }
    if (is_ident_start(c)) {
      size_t j = i;
      while (j < src.size() && is_ident_part(src[j])) ++j;
      out.push_back({Tok::IDENT, src.substr(i, j - i), i});
      i = j;
      continue;
    }
    bool matched = false;
    for (const char* o : ops3)
      if (src.compare(i, 3, o) == 0) {
        out.push_back({Tok::OP, o, i});
        i += 3;
        matched = true;
        break;
      }
    if (matched) continue;
    for (const char* o : ops2)
      if (src.compare(i, 2, o) == 0) {
        out.push_back({Tok::OP, o, i});
        i += 2;
        matched = true;
        break;
      }
    if (matched) continue;
    if (std::strchr("+-*/%^()<>=!,;", c)) {
      out.push_back({Tok::OP, std::string(1, c), i});
      ++i;
      continue;
    }
    throw Error{"unexpected character '" + std::string(1, c) + "' at " + std::to_string(i)};
  }
  out.push_back({Tok::END, "", src.size()});
  return out;
}

int binary_prec(const std::string& op) {
  if (op == "||") return 70;
  if (op == "&&") return 75;
  if (op == "<" || op == ">" || op == "<=" || op == ">=" || op == "==" || op == "!=") return 80;
  if (op == "+" || op == "-") return 95;
  if (op == "*" || op == "/" || op == "%") return 100;
  if (op == "^") return 120;
  return -1;
}

bool is_assign_op(const std::string& op) {
  return op == "=" || op == "+=" || op == "-=" || op == "*=" || op == "/=" || op == "%=" || op == "^=" || op == "&&=" ||
         op == "||=";
}

class Parser {
 public:
  explicit Parser(const std::string& src) : t_(tokenize(src)) {}

  NodeP parse() {
    NodeP n = chain();
    if (peek().kind != Tok::END) fail("unexpected '" + peek().s + "'");
    return n;
  }

 private:
  std::vector<Tok> t_;
  size_t i_ = 0;

  const Tok& peek() const { return t_[i_]; }
  bool at_op(const char* o) const { return peek().kind == Tok::OP && peek().s == o; }
  [[noreturn]] void fail(const std::string& m) const { throw Error{m + " at " + std::to_string(peek().pos)}; }

  static NodeP mk(Node::Kind k) {
    NodeP n(new Node());
    n->kind = k;
    return n;
  }

  bool starts_operand() const {
    const Tok& k = peek();
    if (k.kind == Tok::NUM || k.kind == Tok::IDENT) return true;
    return k.kind == Tok::OP && (k.s == "(" || k.s == "-" || k.s == "!");
  }

  // statements separated by `;` (an empty last statement is Empty)
  NodeP chain() {
    NodeP first = starts_operand() ? tuple() : mk(Node::EMPTY);
    if (!at_op(";")) return first;
    NodeP c = mk(Node::CHAIN);
    c->kids.push_back(std::move(first));
    while (at_op(";")) {
      ++i_;
      c->kids.push_back(starts_operand() ? tuple() : mk(Node::EMPTY));
    }
    return c;
  }

  NodeP tuple() {
    NodeP first = assign();
    if (!at_op(",")) return first;
    NodeP tu = mk(Node::TUPLE);
    tu->kids.push_back(std::move(first));
    while (at_op(",")) {
      ++i_;
      tu->kids.push_back(assign());
    }
    return tu;
  }

  NodeP assign() {
    NodeP lhs = binary(70);
    if (peek().kind == Tok::OP && is_assign_op(peek().s)) {
      if (lhs->kind != Node::VAR) fail("assignment to something that is not a variable");
      NodeP a = mk(Node::ASSIGN);
      a->name = lhs->name;
      std::string op = peek().s;
      ++i_;
      NodeP rhs = assign();  // right-associative
      if (op != "=") {
        // a op= b is a = a op b
        NodeP b = mk(Node::BIN);
        b->name = op.substr(0, op.size() - 1);
        b->kids.push_back(std::move(lhs));
        b->kids.push_back(std::move(rhs));
        rhs = std::move(b);
      }
      a->kids.push_back(std::move(rhs));
      return a;
    }
    return lhs;
  }

  NodeP binary(int min_prec) {
    NodeP lhs = unary();
    for (;;) {
      if (peek().kind != Tok::OP) return lhs;
      const int p = binary_prec(peek().s);
      if (p < 0 || p < min_prec) return lhs;
      NodeP b = mk(Node::BIN);
      b->name = peek().s;
      ++i_;
      NodeP rhs = binary(p + 1);  // left-associative
      b->kids.push_back(std::move(lhs));
      b->kids.push_back(std::move(rhs));
      lhs = std::move(b);
    }
  }

  NodeP unary() {
    if (at_op("-") || at_op("!")) {
      NodeP u = mk(at_op("-") ? Node::NEG : Node::NOT);
      ++i_;
      u->kids.push_back(binary(110));  // only `^` binds tighter
      return u;
    }
    return primary();
  }

  NodeP primary() {
    const Tok k = peek();
    if (k.kind == Tok::NUM) {
      ++i_;
      NodeP n = mk(Node::NUM);
      const bool is_int = k.s.find_first_of(".eE") == std::string::npos;
      if (is_int) {
        errno = 0;
        char* end = nullptr;
        const long long v = std::strtoll(k.s.c_str(), &end, 10);
        if (errno || (v > 9007199254740992LL)) fail("integer literal out of range");
        n->val = sbp::expr::mk((double)v, TAG_INT);
      } else {
        char* end = nullptr;
        const double v = std::strtod(k.s.c_str(), &end);
        if (!end || *end) fail("bad number '" + k.s + "'");
        n->val = sbp::expr::mk(v, TAG_FLOAT);
      }
      return n;
    }
    if (k.kind == Tok::IDENT) {
      ++i_;
      if (k.s == "true" || k.s == "false") {
        NodeP n = mk(Node::BOOL);
        n->val = sbp::expr::mk(k.s == "true" ? 1.0 : 0.0, TAG_BOOL);
        return n;
      }
      if (at_op("(")) {
        ++i_;
        NodeP call = mk(Node::CALL);
        call->name = k.s;
        if (!at_op(")")) {
          NodeP args = chain();
          if (args->kind == Node::TUPLE)
            call->kids = std::move(args->kids);
          else
            call->kids.push_back(std::move(args));
        }
        if (!at_op(")")) fail("expected ')'");
        ++i_;
        return call;
      }
      NodeP v = mk(Node::VAR);
      v->name = k.s;
      return v;
    }
    if (at_op("(")) {
      ++i_;
      NodeP inner = chain();
      if (!at_op(")")) fail("expected ')'");
      ++i_;
      return inner;
    }
    fail(k.kind == Tok::END ? "unexpected end of expression" : "unexpected '" + k.s + "'");
  }

};

const std::map<std::string, int>& functions() {
  static const std::map<std::string, int> f = {
      {"if", FN_IF},
      {"case", FN_CASE},
      {"min", FN_MIN},
      {"max", FN_MAX},
      {"floor", FN_FLOOR},
      {"round", FN_ROUND},
      {"ceil", FN_CEIL},
      {"math::abs", FN_ABS},
      {"math::ln", FN_LN},
      {"math::log", FN_LOG},
      {"math::log2", FN_LOG2},
      {"math::log10", FN_LOG10},
      {"math::exp", FN_EXP},
      {"math::exp2", FN_EXP2},
      {"math::pow", FN_POW},
      {"math::sqrt", FN_SQRT},
      {"math::cbrt", FN_CBRT},
      {"math::hypot", FN_HYPOT},
      {"math::sin", FN_SIN},
      {"math::cos", FN_COS},
      {"math::tan", FN_TAN},
      {"math::asin", FN_ASIN},
      {"math::acos", FN_ACOS},
      {"math::atan", FN_ATAN},
      {"math::atan2", FN_ATAN2},
      {"math::sinh", FN_SINH},
      {"math::cosh", FN_COSH},
      {"math::tanh", FN_TANH},
      {"math::asinh", FN_ASINH},
      {"math::acosh", FN_ACOSH},
      {"math::atanh", FN_ATANH},
      {"math::is_nan", FN_IS_NAN},
      {"math::is_finite", FN_IS_FINITE},
      {"math::is_infinite", FN_IS_INFINITE},
      {"math::is_normal", FN_IS_NORMAL},
  };
  return f;
}

bool arity_ok(int fn, int n) {
  switch (fn) {
    case FN_IF: return n == 3;
    case FN_CASE: return n >= 1 && n % 2 == 1;  // conditions and values, then the default (evalexpr.rs:325-346)
    case FN_MIN:
    case FN_MAX: return n >= 1;
    case FN_LOG:
    case FN_POW:
    case FN_HYPOT:
    case FN_ATAN2: return n == 2;
    default: return n == 1;
  }
}

const char* const kInputNames[IN_COUNT] = {"t", "x", "y", "rho", "rhou", "rhov", "u", "v"};

int input_index(const std::string& s) {
  for (int i = 0; i < IN_COUNT; ++i)
    if (s == kInputNames[i]) return i;
  return -1;
}

class Emitter {
 public:
  Emitter(unsigned inputs, const std::map<std::string, Value>& globals, std::vector<Instr>& code, std::vector<Value>& consts)
      : inputs_(inputs), globals_(globals), code_(code), consts_(consts) {}

  int emit_expression(const Node& root) {
    collect_targets(root);
    const size_t start = code_.size();
    node(root);
    if (depth_ != 1) throw Error{"internal: unbalanced program"};
    return (int)(code_.size() - start);
  }

  const std::map<std::string, int>& slots() const { return slots_; }

 private:
  unsigned inputs_;
  const std::map<std::string, Value>& globals_;
  std::vector<Instr>& code_;
  std::vector<Value>& consts_;
  std::map<std::string, int> slots_;  // assigned so far (in evaluation order)
  std::set<std::string> targets_;     // every name the expression assigns
  int depth_ = 0;

  void collect_targets(const Node& n) {
    if (n.kind == Node::ASSIGN && input_index(n.name) < 0) targets_.insert(n.name);
    for (auto& k : n.kids) collect_targets(*k);
  }

  void emit(int op, int arg, int push) {
    code_.push_back({op, arg});
    depth_ += push;
    if (depth_ > kMaxStack) throw Error{"expression nests too deeply"};
  }

  int constant(const Value& v) {
    for (size_t i = 0; i < consts_.size(); ++i)
      if (consts_[i].tag == v.tag && std::memcmp(&consts_[i].v, &v.v, sizeof(double)) == 0) return (int)i;
    consts_.push_back(v);
    return (int)consts_.size() - 1;
  }

  void read(const std::string& name) {
    const int in = input_index(name);
    if (in >= 0) {
      if (!(inputs_ & (1u << in))) throw Error{"'" + name + "' is not available in this expression"};
      emit(OP_INPUT, in, 1);
      return;
    }
    auto s = slots_.find(name);
    if (s != slots_.end()) {
      emit(OP_LOAD, s->second, 1);
      return;
    }
    if (targets_.count(name)) throw Error{"'" + name + "' is read before the expression assigns it"};
    auto g = globals_.find(name);
    if (g == globals_.end()) throw Error{"unknown variable '" + name + "'"};
    emit(OP_CONST, constant(g->second), 1);
  }

  void node(const Node& n) {
    switch (n.kind) {
      case Node::NUM:
      case Node::BOOL: emit(OP_CONST, constant(n.val), 1); break;
      case Node::EMPTY: emit(OP_EMPTY, 0, 1); break;
      case Node::VAR: read(n.name); break;
      case Node::NEG:
        node(*n.kids[0]);
        emit(OP_NEG, 0, 0);
        break;
      case Node::NOT:
        node(*n.kids[0]);
        emit(OP_NOT, 0, 0);
        break;
      case Node::BIN: {
        static const std::map<std::string, int> ops = {
            {"+", OP_ADD}, {"-", OP_SUB}, {"*", OP_MUL}, {"/", OP_DIV}, {"%", OP_MOD}, {"^", OP_POW}, {"<", OP_LT},
            {">", OP_GT}, {"<=", OP_LE}, {">=", OP_GE}, {"==", OP_EQ}, {"!=", OP_NE}, {"&&", OP_AND}, {"||", OP_OR}};
        node(*n.kids[0]);
        node(*n.kids[1]);
        emit(ops.at(n.name), 0, -1);
        break;
      }
      case Node::ASSIGN: {
        node(*n.kids[0]);
        if (input_index(n.name) >= 0) {
          // the reference stores it, but its reads still see the input
          emit(OP_POP, 0, -1);
          emit(OP_EMPTY, 0, 1);
          break;
        }
        auto s = slots_.find(n.name);
        int slot;
        if (s == slots_.end()) {
          slot = (int)slots_.size();
          if (slot >= kMaxSlots) throw Error{"too many variables"};
          slots_[n.name] = slot;
        } else {
          slot = s->second;
        }
        emit(OP_STORE, slot, 0);
        break;
      }
      case Node::CALL: {
        auto f = functions().find(n.name);
        if (f == functions().end()) throw Error{"unknown function '" + n.name + "'"};
        const int na = (int)n.kids.size();
        if (!arity_ok(f->second, na) || na > kMaxArgs)
          throw Error{"wrong number of arguments (" + std::to_string(na) + ") for '" + n.name + "'"};
        for (auto& k : n.kids) node(*k);
        emit(OP_CALL, f->second | (na << 8), 1 - na);
        break;
      }
      case Node::TUPLE: throw Error{"a tuple is not a value here"};
      case Node::CHAIN:
        for (size_t i = 0; i < n.kids.size(); ++i) {
          node(*n.kids[i]);
          if (i + 1 < n.kids.size()) emit(OP_POP, 0, -1);
        }
        break;
    }
  }
};

// inputs each expression of a form may read (evalexpr.rs:168-208 and :210-290)
unsigned inputs_of(int form, int i) {
  const unsigned base = (1u << IN_T) | (1u << IN_X) | (1u << IN_Y);
  if (i == 0) return base;
  if (i == 1 || i == 2) return base | (1u << IN_RHO);
  return form == FORM_CONSERVATION ? base | (1u << IN_RHO) | (1u << IN_RHOU) | (1u << IN_RHOV)
                                   : base | (1u << IN_RHO) | (1u << IN_U) | (1u << IN_V);
}

void set_err(char* buf, int len, const std::string& m) {
  if (buf && len > 0) std::snprintf(buf, (size_t)len, "%s", m.c_str());
}

}  // namespace

struct sbp_expr_s {
  std::shared_ptr<std::vector<unsigned char>> blob;  // expr::Program + code + constants
};

namespace sbp {
// The program blob of an expression handle (for the runtime).
std::shared_ptr<const std::vector<unsigned char>> expr_blob(sbp_expr e) { return e ? e->blob : nullptr; }
}  // namespace sbp

extern "C" {

int sbp_expr_create(int form, const char* globals, const char* const* exprs, sbp_expr* out, char* errbuf, int errlen) {
  if (!out || !exprs || (form != SBP_EXPR_CONSERVATION && form != SBP_EXPR_PRESSURE)) return SBP_ERR_INVALID;
  for (int i = 0; i < 4; ++i)
    if (!exprs[i]) return SBP_ERR_INVALID;
  double gamma;
  if (sbp_get_gamma(&gamma) != SBP_OK) return SBP_ERR_GAMMA;  // default_context reads GAMMA (evalexpr.rs:298-301)
  try {
    // the default context: math constants and GAMMA, then the globals
    std::map<std::string, Value> g = {{"PI", mk(M_PI, TAG_FLOAT)},
                                      {"E", mk(M_E, TAG_FLOAT)},
                                      {"TAU", mk(2.0 * M_PI, TAG_FLOAT)},
                                      {"GAMMA", mk(gamma, TAG_FLOAT)}};
    if (globals && *globals) {
      std::vector<Instr> code;
      std::vector<Value> consts;
      Parser p(globals);
      NodeP root = p.parse();
      Emitter em(0u, g, code, consts);
      em.emit_expression(*root);
      Value r, slots[kMaxSlots];
      const double none[IN_COUNT] = {0, 0, 0, 0, 0, 0, 0, 0};
      if (!run(code.data(), (int)code.size(), consts.data(), none, r, slots))
        throw Error{"the globals do not evaluate"};
      for (auto& kv : em.slots()) g[kv.first] = slots[kv.second];
    }
    std::vector<Instr> code;
    std::vector<Value> consts;
    Program hdr{};
    hdr.form = form;
    hdr.gamma = gamma;
    for (int i = 0; i < 4; ++i) {
      Parser p(exprs[i]);
      NodeP root = p.parse();
      Emitter em(inputs_of(form, i), g, code, consts);
      hdr.off[i] = (int)code.size();
      hdr.len[i] = em.emit_expression(*root);
    }
    hdr.ncode = (int)code.size();
    hdr.nconst = (int)consts.size();
    auto blob = std::make_shared<std::vector<unsigned char>>(program_bytes(hdr.ncode, hdr.nconst), 0);
    std::memcpy(blob->data(), &hdr, sizeof(hdr));
    std::memcpy(blob->data() + sizeof(hdr), code.data(), sizeof(Instr) * code.size());
    Program* P = reinterpret_cast<Program*>(blob->data());
    std::memcpy(const_cast<Value*>(consts_of(P)), consts.data(), sizeof(Value) * consts.size());
    auto* e = new sbp_expr_s();
    e->blob = blob;
    *out = e;
    return SBP_OK;
  } catch (const Error& err) {
    set_err(errbuf, errlen, err.msg);
    return SBP_ERR_EXPR;
  } catch (const std::exception& err) {
    set_err(errbuf, errlen, err.what());
    return SBP_ERR_EXPR;
  }
}

int sbp_expr_destroy(sbp_expr e) {
  if (!e) return SBP_ERR_INVALID;
  delete e;
  return SBP_OK;
}

int sbp_expr_eval_host(sbp_expr e, double t, int64_t n, const double* x, const double* y, double* out) {
  if (!e || n < 0 || (n > 0 && (!x || !y || !out))) return SBP_ERR_INVALID;
  const Program* P = reinterpret_cast<const Program*>(e->blob->data());
  for (int64_t i = 0; i < n; ++i) {
    double q[4];
    if (!eval_point(P, t, x[i], y[i], q)) return SBP_ERR_EXPR;
    for (int c = 0; c < 4; ++c) out[c * n + i] = q[c];
  }
  return SBP_OK;
}

}  // extern "C"
