"""Source-to-source translator for OPS applications (MI355X backend).

Reads C++ application files written against the OPS dev API
(ops_par_loop(kernel, name, block, dim, range, args...), user kernels taking
ACC<T> references, ops/c/include/ops_seq_v2.h:327-335) and writes, per
application file app.cpp:

  app_ops.cpp      the application with every ops_par_loop(kernel, ...) call
                   rewritten to ops_par_loop_<kernel>(...) (the reference
                   translator's rewrite, ops_translator/ops-translator
                   "loop_host" stubs: resources/templates/cpp/loop_host.cpp.j2:12-33),
                   its OPS include pointed at ops_mi355x.h and the kernel
                   definitions and kernel-header includes dropped;
  app_opsmi_kernels.h    declarations of the generated stubs;
  app_opsmi_kernels.hip  one HIP kernel per user kernel plus its host stub,
                         built on include/ops_mi355x_device.h and
                         include/ops_mi355x_gen.h.

The user kernel's body is emitted unchanged as the call operator of a
struct whose members are the ops_decl_const constants the body names, so the
body's arithmetic is the user's own (compile with -ffp-contract=off for
the reference's sequential rounding).  Requirements, as for the reference
translator: the kernel is a literal identifier, the loop's dim is an integer
literal, every ops_arg_* call spells its type as a string literal and its
access as an OPS_* name, ops_arg_reduce dims are integer literals (an
ops_arg_gbl whose dim is not a literal gets a copy of up to 64 values, and
a call passing more is refused at run time), and the kernel is defined in
the application file or a header it includes with quotes.

Helper functions the kernels call live in the application's own headers,
declared with the OPS_FUN_PREFIX macro as for the reference's generated code
(e.g. apps/c/poisson/user_types.h); the generated kernel file defines it as
__host__ __device__ and includes the application's local headers other than
the kernel headers.

Usage: python -m ops_amd.translator.opsmi_translate app.cpp [...] -o outdir [-I dir]
                                                   [--shared LIB.so | --exe PROGRAM]
"""
import argparse
import os
import re
import sys

TYPES = {"double": ("OPSMI_GEN_F64", 8), "float": ("OPSMI_GEN_F32", 4), "int": ("OPSMI_GEN_I32", 4),
         "half": ("OPSMI_GEN_F16", 2)}
GBL_CAP = 64  # values a generated kernel holds of an ops_arg_gbl whose dim is not a literal
ACCESS = {"OPS_READ", "OPS_WRITE", "OPS_RW", "OPS_INC", "OPS_MIN", "OPS_MAX"}


class TranslateError(Exception):
    pass


# ------------------------------------------------------------------ lexing
def mask(src):
    """src with comments and the insides of string/char literals replaced by
    spaces (newlines kept), so offsets match the original text."""
    out = list(src)
    i, n = 0, len(src)
    while i < n:
        c = src[i]
        if src.startswith("//", i):
            j = src.find("\n", i)
            j = n if j < 0 else j
            for k in range(i, j):
                out[k] = " "
            i = j
        elif src.startswith("/*", i):
            j = src.find("*/", i + 2)
            j = n if j < 0 else j + 2
            for k in range(i, j):
                if out[k] != "\n":
                    out[k] = " "
            i = j
        elif c in "\"'":
            j = i + 1
            while j < n and src[j] != c:
                j += 2 if src[j] == "\\" else 1
            for k in range(i + 1, min(j, n)):
                if out[k] != "\n":
                    out[k] = " "
            i = j + 1
        else:
            i += 1
    return "".join(out)


def match_close(m, i):
    """index of the bracket closing the one at m[i]"""
    pairs = {"(": ")", "[": "]", "{": "}"}
    stack = []
    for j in range(i, len(m)):
        c = m[j]
        if c in pairs:
            stack.append(pairs[c])
        elif c in ")]}":
            if not stack or stack.pop() != c:
                raise TranslateError(f"unbalanced '{c}' at offset {j}")
            if not stack:
                return j
    raise TranslateError(f"no closing bracket for offset {i}")


def split_top(m, a, b):
    """(start, end) spans of the comma-separated items of m[a:b]"""
    spans, depth, s = [], 0, a
    for j in range(a, b):
        c = m[j]
        if c in "([{":
            depth += 1
        elif c in ")]}":
            depth -= 1
        elif c == "," and depth == 0:
            spans.append((s, j))
            s = j + 1
    if m[s:b].strip():
        spans.append((s, b))
    return spans


def strip_span(src, span):
    a, b = span
    while a < b and src[a].isspace():
        a += 1
    while b > a and src[b - 1].isspace():
        b -= 1
    return a, b


def string_literal(text, what):
    t = text.strip()
    if not (len(t) >= 2 and t[0] == '"' and t[-1] == '"'):
        raise TranslateError(f"{what} must be a string literal, got {t!r}")
    return t[1:-1]


def int_literal(text):
    t = text.strip()
    return int(t) if re.fullmatch(r"\d+", t) else None


# ------------------------------------------------------------------ model
class Arg:
    def __init__(self, kind, typ=None, acc=None, dim=None, multi=False, expr=None):
        self.kind = kind      # dat | gbl | reduce | idx
        self.typ = typ        # double | float | int
        self.acc = acc        # OPS_* access
        self.dim = dim        # gbl/reduce component count
        self.multi = multi    # dat: may have more than one component
        self.expr = expr      # dat: the dat expression as written (whitespace removed)


class Loop:
    def __init__(self, kernel, ndim, args, span, name_span):
        self.kernel, self.ndim, self.args = kernel, ndim, args
        self.span = span            # (start of 'ops_par_loop', end of the first argument's comma)
        self.name_span = name_span
        self.path = None            # the file the call is in


class Kernel:
    def __init__(self, name, params, body, path, span):
        self.name, self.params, self.body, self.path, self.span = name, params, body, path, span


def parse_arg(src, m, span, where):
    a, b = strip_span(src, span)
    text_m = m[a:b]
    mm = re.match(r"(ops_arg_dat_opt|ops_arg_dat|ops_arg_gbl|ops_arg_reduce|ops_arg_idx)\s*(<[^>]*>)?\s*\(", text_m)
    if not mm:
        raise TranslateError(f"{where}: argument {src[a:b]!r} is not an ops_arg_* call")
    fn = mm.group(1)
    op = a + mm.end() - 1
    cl = match_close(m, op)
    items = [src[s:e].strip() for s, e in split_top(m, op + 1, cl)]
    if fn == "ops_arg_idx":
        return Arg("idx")
    if fn in ("ops_arg_dat", "ops_arg_dat_opt"):
        if len(items) < 5:
            raise TranslateError(f"{where}: {fn} needs 5 arguments")
        typ = string_literal(items[3], f"{where}: {fn} type")
        acc = items[4]
        d = int_literal(items[1])
        return Arg("dat", typ, acc, d, multi=(d != 1), expr=re.sub(r"\s+", "", items[0]))
    if len(items) != 4:
        raise TranslateError(f"{where}: {fn} needs 4 arguments")
    typ = string_literal(items[2], f"{where}: {fn} type")
    acc = items[3]
    d = int_literal(items[1])
    kind = "gbl" if fn == "ops_arg_gbl" else "reduce"
    if kind == "gbl" and acc != "OPS_READ":
        kind = "reduce"
    if d is None:
        if kind == "reduce":
            raise TranslateError(f"{where}: {fn} dim must be an integer literal for a reduction")
        d = GBL_CAP  # a run-time count: the kernel's copy holds up to GBL_CAP values
    return Arg(kind, typ, acc, d)


def find_loops(src, m, path):
    loops = []
    for mo in re.finditer(r"\bops_par_loop\s*\(", m):
        op = mo.end() - 1
        cl = match_close(m, op)
        spans = split_top(m, op + 1, cl)
        where = f"{path}:{src.count(chr(10), 0, mo.start()) + 1}"
        if len(spans) < 5:
            raise TranslateError(f"{where}: ops_par_loop needs kernel, name, block, dim, range")
        kernel = src[slice(*strip_span(src, spans[0]))]
        if not re.fullmatch(r"[A-Za-z_]\w*", kernel):
            raise TranslateError(f"{where}: kernel {kernel!r} is not an identifier")
        nd = int_literal(src[slice(*spans[3])])
        if nd not in (1, 2, 3):
            raise TranslateError(f"{where}: loop dim must be the integer literal 1, 2 or 3")
        args = [parse_arg(src, m, sp, where) for sp in spans[5:]]
        for x in args:
            if x.kind != "idx":
                if x.typ not in TYPES:
                    raise TranslateError(f"{where}: type {x.typ!r} (only double, float, int, half)")
                if x.acc not in ACCESS:
                    raise TranslateError(f"{where}: access {x.acc!r} is not an OPS_* access")
        loops.append(Loop(kernel, nd, args, (mo.start(), spans[0][1] + 1), spans[1]))
        loops[-1].path = path
    return loops


def loop_body_pairs(src, m, loops):
    """(last, first) ops_par_loop calls of every while/for/do body holding
    two or more of them: the pair that meets in the queue when the body
    runs again (the last loop of one pass, then the first of the next)."""
    out = []
    for mo in re.finditer(r"\b(while|for)\s*\(|\bdo\s*\{", m):
        if mo.group(0).startswith("do"):
            ob = mo.end() - 1
        else:
            cl = match_close(m, mo.end() - 1)
            j = cl + 1
            while j < len(m) and m[j].isspace():
                j += 1
            if j >= len(m) or m[j] != "{":
                continue
            ob = j
        cb = match_close(m, ob)
        inside = [lp for lp in loops if ob < lp.span[0] < cb]
        if len(inside) >= 2:
            out.append((inside[-1], inside[0]))
    return out


def find_consts(src, m):
    """ops_decl_const("name", dim, "type", ptr) calls: name -> (dim, type);
    dim is None when it is not an integer literal (the kernels then hold up
    to OPSMI_CONST_BYTES bytes of it), type a basic type or the name of a
    user type (a struct the application's headers define)."""
    out = {}
    for mo in re.finditer(r"\bops_decl_const\s*(<[^>]*>)?\s*\(", m):
        op = mo.end() - 1
        cl = match_close(m, op)
        items = [src[s:e].strip() for s, e in split_top(m, op + 1, cl)]
        if len(items) != 4:
            continue
        try:
            name = string_literal(items[0], "constant name")
            typ = string_literal(items[2], "constant type")
        except TranslateError:
            continue
        d = int_literal(items[1])
        if not re.fullmatch(r"[A-Za-z_]\w*", typ) or (d is not None and d < 1):
            raise TranslateError(f"ops_decl_const({items[0]}): type must be a type name and dim a positive count")
        out[name] = (d, typ)
    return out


# (static, const and constexpr variables have internal linkage: no clash)
GLOBAL_DEF = re.compile(r"^\s*(?!extern\b|typedef\b|using\b|return\b|namespace\b|template\b|static\b|const\b"
                        r"|constexpr\b)(?:unsigned\s+|signed\s+|long\s+|short\s+)*"
                        r"[A-Za-z_][\w:]*(?:\s*<[^;()]*>)?[\s\*&]+[A-Za-z_]\w*\s*(?:\[[^\]]*\]\s*)*"
                        r"(?:=[^;]*)?(?:,\s*\**\s*[A-Za-z_]\w*\s*(?:\[[^\]]*\]\s*)*(?:=[^;]*)?)*;")


def defines_globals(src):
    """True when a header defines variables at file scope (not extern
    declarations): the application's main file includes such a header once,
    and a generated kernels file must not define them again."""
    m = mask(src)
    depth, start = 0, 0
    for i, ch in enumerate(m):
        if ch == "{":
            depth += 1
        elif ch == "}":
            depth -= 1
            start = i + 1
        elif ch == ";" and depth == 0:
            stmt = m[start:i + 1]
            start = i + 1
            # drop preprocessor lines
            stmt = "\n".join(ln for ln in stmt.splitlines() if not ln.lstrip().startswith("#"))
            if "(" not in stmt and GLOBAL_DEF.match(stmt.strip().replace("\n", " ")):
                return True
        elif ch == "\n" and depth == 0 and m[start:i].lstrip().startswith("#"):
            start = i + 1
    return False


def defines_functions(src):
    """True when a file defines functions with external linkage at file
    scope (not inline, static, constexpr or templates), as an implementation
    file the application includes (CloverLeaf's profile.cpp) does: a
    generated kernels file that included it too would define them twice."""
    m = mask(src)
    depth, start = 0, 0
    for i, ch in enumerate(m):
        if ch == "{":
            if depth == 0:
                head = "\n".join(ln for ln in m[start:i].splitlines() if not ln.lstrip().startswith("#")).strip()
                if re.search(r"\)\s*(?:const\s*)?(?:noexcept\s*)?$", head) and \
                        not re.search(r"\b(?:inline|static|constexpr|template|struct|class|namespace|extern\s*\"C\")\b",
                                      head) and not re.match(r"^(?:if|for|while|switch)\b", head):
                    return True
            depth += 1
        elif ch == "}":
            depth -= 1
            if depth == 0:
                start = i + 1
        elif ch == ";" and depth == 0:
            start = i + 1
        elif ch == "\n" and depth == 0 and m[start:i].lstrip().startswith("#"):
            start = i + 1
    return False


def const_member(c, d, t):
    """the member holding constant c in a kernel's functor"""
    if d is None:
        return f"{t} {c}[OPSMI_CONST_CAP({t})];"
    return f"{t} {c}{'[' + str(d) + ']' if d > 1 else ''};"


def const_fetch(c, d, t):
    """the stub's copy of constant c's current value into the functor"""
    if d is not None and t in TYPES:
        return f"ops_mi_gen_const(\"{c}\", \"{t}\", {d}, &b.k.{c});"
    cap = f"(int)OPSMI_CONST_CAP({t})" if d is None else str(d)
    return f"ops_mi_gen_const_sized(\"{c}\", \"{t}\", sizeof({t}), {cap}, {'1' if d is not None else '0'}, &b.k.{c});"


class Program:
    """What the source files of one program share (the reference translator
    sees the whole program too): constants declared in any file, kernels
    defined in any file or local header, and the headers that define the
    kernels the program's loops call (dropped from every translated file)."""

    def __init__(self, apps, inc_dirs=()):
        self.files, seen = [], set()
        self.consts, names = {}, set()
        for app in apps:
            src = open(app).read()
            m = mask(src)
            for p, s2, m2 in [(app, src, m)] + [(p, None, None) for _, _, p, _ in
                                                  local_includes(app, src, m, list(inc_dirs))]:
                if p in seen:
                    continue
                seen.add(p)
                if s2 is None:
                    s2 = open(p).read()
                    m2 = mask(s2)
                self.files.append((p, s2, m2))
                self.consts.update(find_consts(s2, m2))
            names |= {lp.kernel for lp in find_loops(src, m, app)}
        self.kernel_headers = set()
        for k in names:
            path = find_kernel(k, self.files).path
            if path not in apps:
                self.kernel_headers.add(path)


def find_kernel(name, files):
    """(Kernel) definition `void name(params) { body }` in one of files"""
    for path, src, m in files:
        for mo in re.finditer(r"\bvoid\s+" + re.escape(name) + r"\s*\(", m):
            op = mo.end() - 1
            cl = match_close(m, op)
            j = cl + 1
            while j < len(m) and m[j].isspace():
                j += 1
            if j >= len(m) or m[j] != "{":
                continue  # a declaration
            bcl = match_close(m, j)
            params = []
            for s, e in split_top(m, op + 1, cl):
                # the parameter as written, without its comments
                text = " ".join(re.sub(r"//[^\n]*|/\*.*?\*/", " ", src[s:e], flags=re.S).split())
                pm = re.search(r"([A-Za-z_]\w*)\s*$", m[s:e].strip())
                if not pm:
                    raise TranslateError(f"{path}: kernel {name}: cannot read parameter {text!r}")
                params.append((text, pm.group(1)))
            start = mo.start()
            # a 'static' or 'inline' in front belongs to the definition
            pre = re.search(r"((?:static|inline)\s+)+$", m[max(0, start - 32):start])
            if pre:
                start -= len(pre.group(0))
            return Kernel(name, params, src[j:bcl + 1], path, (start, bcl + 1))
    raise TranslateError(f"kernel {name} is not defined in {', '.join(p for p, _, _ in files)}")


def local_includes(path, src, m, inc_dirs):
    out = []
    for mo in re.finditer(r'#\s*include\s*"([^"]+)"', src):
        if m[mo.start()] != "#":
            continue  # inside a comment
        for d in [os.path.dirname(os.path.abspath(path))] + inc_dirs:
            p = os.path.join(d, mo.group(1))
            if os.path.isfile(p):
                out.append((mo.start(), mo.end(), os.path.abspath(p), mo.group(1)))
                break
    return out


# ------------------------------------------------------------------ emit
def accessed_points(body, pname, ndim):
    """(component, dx, dy, dz) of every ACC call pname(...) in body whose
    indices are integer literals; None when the kernel touches the accessor
    any other way (non-literal indices, combine_* calls, passing it to a
    helper, binding another name to it) or never names it, since then the
    points it reaches are unknown (the stencil is not checked against the
    kernel, and the loop is not fused on the assumption that the dat is
    touched only at the point)."""
    bm = mask(body)
    pts = set()
    names = list(re.finditer(r"\b" + re.escape(pname) + r"\b", bm))
    calls = list(re.finditer(r"\b" + re.escape(pname) + r"\s*\(", bm))
    if not names or len(names) != len(calls):
        return None
    for mo in calls:
        if address_taken(bm, mo.start()):
            return None
        # skip member calls (pname.combine_max(...) has no '(' right after the name)
        op = mo.end() - 1
        cl = match_close(bm, op)
        items = [bm[s:e].strip() for s, e in split_top(bm, op + 1, cl)]
        vals = []
        for it in items:
            if not re.fullmatch(r"[+-]?\s*\d+", it):
                return None
            vals.append(int(it.replace(" ", "")))
        if len(vals) == ndim:
            c, off = -1, vals
        elif len(vals) == ndim + 1:
            c, off = vals[0], vals[1:]
        else:
            return None
        off = off + [0] * (3 - len(off))
        pts.add((c, off[0], off[1], off[2]))
    return sorted(pts)


def address_taken(bm, pos):
    """True when the name at pos is the operand of a unary & (its address
    taken: pointer arithmetic from it reaches points no index lists)"""
    j = pos - 1
    while j >= 0 and bm[j].isspace():
        j -= 1
    return j >= 0 and bm[j] == "&" and (j == 0 or bm[j - 1] != "&")


def legacy_points(body, pname, ndim):
    """accessed_points for a legacy kernel: every use of pname must be an
    index pname[OPS_ACCn(dx, ...)] or pname[OPS_ACC_MDn(d, dx, ...)] with
    integer literals; None otherwise."""
    bm = mask(body)
    pts = set()
    names = list(re.finditer(r"\b" + re.escape(pname) + r"\b", bm))
    uses = list(re.finditer(r"\b" + re.escape(pname) + r"\s*\[\s*OPS_ACC(_MD)?\d+\s*\(", bm))
    if not names or len(names) != len(uses):
        return None
    for mo in uses:
        if address_taken(bm, mo.start()):
            return None
        op = mo.end() - 1
        cl = match_close(bm, op)
        rest = bm[cl + 1:].lstrip()
        if not rest.startswith("]"):
            return None
        items = [bm[a:b].strip() for a, b in split_top(bm, op + 1, cl)]
        vals = []
        for it in items:
            if not re.fullmatch(r"[+-]?\s*\d+", it):
                return None
            vals.append(int(it.replace(" ", "")))
        if mo.group(1):
            if len(vals) < 2:
                return None
            c, off = vals[0], vals[1:]
        else:
            c, off = -1, vals
        if len(off) > ndim:
            return None
        off = off + [0] * (3 - len(off))
        pts.add((c, off[0], off[1], off[2]))
    return sorted(pts)


LEGACY_PARAM = re.compile(r"^(?:const\s+)?([A-Za-z_][\w:]*)\s*(?:const\s*)?\*\s*(?:(?:__restrict__|restrict|__restrict)\s+)?"
                          r"([A-Za-z_]\w*)$")


def is_legacy(kern, loop):
    """True when the kernel takes its dat arguments as T * pointers indexed
    with the OPS_ACCn / OPS_ACC_MDn macros (the legacy API,
    ops/c/include/ops_macros.h), False when they are ACC<T> references."""
    forms = set()
    for (text, pname), a in zip(kern.params, loop.args):
        if a.kind != "dat":
            continue
        if re.search(r"\bACC\s*<", text):
            forms.add("acc")
        elif LEGACY_PARAM.match(text.strip()) and re.search(r"\bOPS_ACC(?:_MD)?\d+\s*\(", kern.body):
            forms.add("legacy")
        else:
            raise TranslateError(f"kernel {kern.name}: parameter {text!r} for an ops_arg_dat must be an ACC<T> "
                                 "reference (or a T * pointer indexed with OPS_ACCn)")
    if len(forms) > 1:
        raise TranslateError(f"kernel {kern.name}: mixes ACC<T> and pointer (OPS_ACCn) dat parameters")
    return forms == {"legacy"}


def kernel_code(loop, kern, consts):
    K = loop.kernel
    n = len(loop.args)
    if len(kern.params) != n:
        raise TranslateError(f"kernel {K} takes {len(kern.params)} parameters, the loop passes {n} arguments")
    used = [c for c in sorted(consts) if re.search(r"\b" + re.escape(c) + r"\b", mask(kern.body))]
    legacy = is_legacy(kern, loop)
    listing = legacy_points if legacy else accessed_points
    points = [listing(kern.body, pname, loop.ndim) if a.kind == "dat" else []
              for (_, pname), a in zip(kern.params, loop.args)]
    L = []
    L.append(f"// ---------------------------------------------------------------- {K}")
    L.append(f"struct opsk_{K} {{")
    for c in used:
        d, t = consts[c]
        L.append("  " + const_member(c, d, t))
    # every ACC<T> parameter becomes a template type: the plain kernels pass
    # ::opsmi_gen::Acc accessors (restrict-qualified or general), fused
    # kernels also pass accessors that capture a write or hold values
    # another loop computed
    ps, tps = [], []
    for i, ((text, pname), a) in enumerate(zip(kern.params, loop.args)):
        if a.kind == "dat":
            if legacy:
                # T *p indexed with OPS_ACCn(...): an accessor the offsets resolve against
                new, cnt = f"const OPSMI_A{i} &{pname}", 1
            else:
                new, cnt = re.subn(r"\bACC\s*<\s*([^>]+?)\s*>", f"OPSMI_A{i}", text)
            if cnt != 1:
                raise TranslateError(f"kernel {K}: parameter {text!r} for an ops_arg_dat must be an ACC<T> reference")
            ps.append(new)
            tps.append(f"class OPSMI_A{i}")
        else:
            ps.append(text)
    tmpl = f"template <{', '.join(tps)}> " if tps else ""
    L.append(f"  {tmpl}__device__ __forceinline__ void operator()({', '.join(ps)}) const {kern.body}")
    L.append("};")
    # by-value data: constants and gbl values
    L.append(f"struct opsb_{K} {{")
    L.append(f"  opsk_{K} k;")
    for i, a in enumerate(loop.args):
        if a.kind == "gbl":
            L.append(f"  {a.typ} g{i}[{a.dim}];")
    L.append("};")
    L.append(f"struct opsa_{K} {{")
    L.append("  ::opsmi_gen::Loop L;")
    for i, a in enumerate(loop.args):
        if a.kind == "dat":
            L.append(f"  opsmi_gen_dat d{i};")
        elif a.kind == "reduce":
            L.append(f"  {a.typ} *r{i};")
    L.append(f"  opsb_{K} b;")
    L.append("};")
    # two kernels: dat pointers restrict-qualified (OPS writes a dat only at
    # the point itself, so a thread's points are independent), and the
    # general one with plain pointers and x steps from the launch, for loops
    # that pass one dat twice or whose stencils do not advance in a
    # dimension (dats of size 1 there, several points on one element)
    dats = [i for i, a in enumerate(loop.args) if a.kind == "dat"]
    for suffix, qual, gen in (("", "__restrict__ ", "false"), ("_alias", "", "true")):
        ptrs = "".join(f", {'const ' if loop.args[i].acc == 'OPS_READ' else ''}{loop.args[i].typ} *{qual}p{i}"
                       for i in dats)
        L.append(f"__global__ __launch_bounds__(256) void opsg_{K}{suffix}(const opsa_{K} A{ptrs}) {{")
        L.extend(red_init_lines(loop, ""))
        # adjacent points per lane where the accesses merge (restrict) and no
        # reduction's order depends on the layout
        adj = "true" if qual and not any(a.kind == "reduce" for a in loop.args) else "false"
        L.append(f"  ::opsmi_gen::for_each_point<{adj}>(A.L, [&](int x, int y, int z) {{")
        L.extend(point_lines(loop, "A", "", gen))
        L.append("  });")
        L.extend(red_store_lines(loop, "A", ""))
        L.append("}")
    L.append(f"void opsfill_{K}(opsa_{K} &A, const opsmi_gen_launch *L, const void *blob) {{")
    L.append("  A.L = ::opsmi_gen::loop_of(L);")
    for i, a in enumerate(loop.args):
        if a.kind == "dat":
            L.append(f"  A.d{i} = L->dat[{i}];")
        elif a.kind == "reduce":
            L.append(f"  A.r{i} = ({a.typ} *)L->partials[{i}];")
    L.append("  memcpy(&A.b, blob, sizeof(A.b));")
    L.append("}")
    L.append(f"void opsl_{K}(const opsmi_gen_launch *L, const void *blob, void *stream) {{")
    L.append(f"  opsa_{K} A;")
    L.append(f"  opsfill_{K}(A, L, blob);")
    ptrs = launch_ptrs(loop, "L")
    L.append(f"  auto kern = L->alias ? opsg_{K}_alias : opsg_{K};")
    L.append(f"  hipLaunchKernelGGL(kern, dim3(L->grid[0], L->grid[1], L->grid[2]), "
             f"dim3(::opsmi_gen::kWave, ::opsmi_gen::kRows), 0, (hipStream_t)stream, A{ptrs});")
    L.append("}")
    return "\n".join(L), used, points


def red_init_lines(loop, pfx):
    L = []
    for i, a in enumerate(loop.args):
        if a.kind == "reduce":
            L.append(f"  {a.typ} {pfx}red{i}[{a.dim}];")
            L.append(f"  for (int q = 0; q < {a.dim}; q++) {pfx}red{i}[q] = ::opsmi_gen::ident<{a.typ}, {a.acc}>();")
    return L


def point_lines(loop, S, pfx, gen="false"):
    """Accessors of loop's arguments at point (x, y, z) over pointers
    {pfx}p<i>, then the call of its kernel (member k of {S}.b)."""
    L, call = [], []
    for i, a in enumerate(loop.args):
        if a.kind == "dat":
            L.append(f"    ::opsmi_gen::Acc<{a.typ}, {loop.ndim}, {'true' if a.multi else 'false'}, {gen}> "
                     f"{pfx}a{i}({pfx}p{i}, {S}.d{i}, x, y, z);")
            call.append(f"{pfx}a{i}")
        elif a.kind == "idx":
            L.append(f"    int {pfx}idx{i}[3] = {{x, y, z}};")
            call.append(f"{pfx}idx{i}")
        elif a.kind == "gbl":
            call.append(f"{S}.b.g{i}")
        else:
            call.append(f"{pfx}red{i}")
    L.append(f"    {S}.b.k({', '.join(call)});")
    return L


def red_store_lines(loop, S, pfx):
    return [f"  ::opsmi_gen::store_partials<{a.typ}, {a.acc}, {a.dim}>({pfx}red{i}, {S}.r{i});"
            for i, a in enumerate(loop.args) if a.kind == "reduce"]


def launch_ptrs(loop, Lname):
    return "".join(f", ({a.typ} *){Lname}->dat[{i}].base" for i, a in enumerate(loop.args) if a.kind == "dat")


def chain_code(ls, calls):
    """One kernel running the kernels of the loops ls (two or more) in turn
    at every point, and its launch function (registered for the chain; the
    library decides per call whether the queued loops may run this way).
    calls are the run's ops_par_loop calls: dat arguments written with the
    same expression share one restrict-qualified pointer (the library runs
    the kernel only when those name one dat and the others distinct dats).
    None when two arguments with one expression differ in element type."""
    F = "__".join(lp.kernel for lp in ls)
    groups, gtyp, gconst, flat = {}, [], [], []
    for k, (lp, c) in enumerate(zip(ls, calls)):
        for i, a in enumerate(lp.args):
            if a.kind != "dat":
                flat.append(-1)
                continue
            e = c.args[i].expr
            if e not in groups:
                groups[e] = len(gtyp)
                gtyp.append(a.typ)
                gconst.append(True)
            j = groups[e]
            if gtyp[j] != a.typ:
                return None
            gconst[j] = gconst[j] and a.acc == "OPS_READ"
            flat.append(j)
    adj = "false" if any(a.kind == "reduce" for lp in ls for a in lp.args) else "true"
    L = [f"// ---------------------------------------------------------------- {' then '.join(lp.kernel for lp in ls)}",
         "// (one restrict-qualified pointer per distinct dat)",
         f"__global__ __launch_bounds__(256) void opsg_chain_{F}("
         + ", ".join(f"const opsa_{lp.kernel} A{k}" for k, lp in enumerate(ls))
         + "".join(f", {'const ' if gconst[j] else ''}{t} *__restrict__ g{j}" for j, t in enumerate(gtyp))
         + ") {"]
    pos = 0
    for k, lp in enumerate(ls):
        for i, a in enumerate(lp.args):
            if a.kind == "dat":
                L.append(f"  {'const ' if a.acc == 'OPS_READ' else ''}{a.typ} *q{k}_p{i} = g{flat[pos]};")
            pos += 1
    for k, lp in enumerate(ls):
        L += red_init_lines(lp, f"q{k}_")
    L.append(f"  ::opsmi_gen::for_each_point<{adj}>(A0.L, [&](int x, int y, int z) {{")
    for k, lp in enumerate(ls):
        L += point_lines(lp, f"A{k}", f"q{k}_")
    L.append("  });")
    for k, lp in enumerate(ls):
        L += red_store_lines(lp, f"A{k}", f"q{k}_")
    L.append("}")
    first = {}
    pos = 0
    for k, lp in enumerate(ls):
        for i, a in enumerate(lp.args):
            if flat[pos] >= 0:
                first.setdefault(flat[pos], (k, i))
            pos += 1
    L.append(f"void opsc_{F}(const opsmi_gen_launch *const *L, const void *const *blob, void *stream) {{")
    for k, lp in enumerate(ls):
        L.append(f"  opsa_{lp.kernel} A{k};")
        L.append(f"  opsfill_{lp.kernel}(A{k}, L[{k}], blob[{k}]);")
    L.append(f"  hipLaunchKernelGGL(opsg_chain_{F}, dim3(L[0]->grid[0], L[0]->grid[1], L[0]->grid[2]), "
             f"dim3(::opsmi_gen::kWave, ::opsmi_gen::kRows), 0, (hipStream_t)stream, "
             + ", ".join(f"A{k}" for k in range(len(ls)))
             + "".join(f", ({gtyp[j]} *)L[{first[j][0]}]->dat[{first[j][1]}].base" for j in range(len(gtyp)))
             + ");")
    L.append("}")
    fns = ", ".join(f"&opsl_{lp.kernel}" for lp in ls)
    return "\n".join(L), (f"{{ static const opsmi_gen_launch_fn c[] = {{{fns}}}; "
                          f"static const int g[] = {{{', '.join(map(str, flat))}}}; "
                          f"ops_mi_gen_fused_chain(c, {len(ls)}, g, {len(flat)}, &opsc_{F}); }}")


def nb_links(l1, l2, pts2):
    """(i, j): argument i of l2 reads, at stencil offsets other than the
    point itself, the dat l1 writes through argument j (same dat
    expression, same type and component count; l1 writes it with OPS_WRITE,
    l2 reads it with OPS_READ at literal offsets).  Empty when no argument
    of l2 reads l1's output at a neighbour."""
    links, nonzero = [], False
    for i, b in enumerate(l2.args):
        if b.kind != "dat" or b.acc != "OPS_READ" or pts2[i] is None:
            continue
        for j, a in enumerate(l1.args):
            if a.kind == "dat" and a.acc == "OPS_WRITE" and a.expr == b.expr and a.typ == b.typ and \
               a.multi == b.multi and (not a.multi or a.dim is not None):
                links.append((i, j))
                nonzero = nonzero or any(q[1:] != (0, 0, 0) for q in pts2[i])
                break
    return links if nonzero else []


def nb_fused_code(l1, l2, pts1, pts2, links):
    """One kernel running, at each point p, the first loop's kernel and then
    the second's, where the second reads at neighbours p + o what the first
    writes (the linked arguments): the first loop's output at p + o is
    recomputed there into registers (its kernel run at p + o with its written
    arguments captured, its reductions dropped) when p + o lies in the range,
    and read from memory (not written by the first loop) when it does not.
    The library launches it only when that gives the results of the two
    loops run one after the other (genloop.hip)."""
    K1, K2 = l1.kernel, l2.kernel
    F = f"{K1}__{K2}"
    nd = l1.ndim
    lk = dict(links)
    ptrs = []
    for pfx, lp in (("f", l1), ("s", l2)):
        ptrs += [f"{'const ' if a.acc == 'OPS_READ' else ''}{a.typ} *{pfx}p{i}"
                 for i, a in enumerate(lp.args) if a.kind == "dat"]
    L = [f"// ---------------------------------------------------------------- {K2} reading {K1}'s output at neighbours"]
    # the held values of each linked argument: one slot per listed (component, offset)
    slots = {}
    for i, j in links:
        b = l2.args[i]
        sl = pts2[i]
        slots[i] = sl
        L.append(f"struct opssh_{F}_{i} {{")
        L.append(f"  {b.typ} v[{len(sl)}];")
        if b.multi:
            sig = ", ".join(["int c"] + [f"int d{q}" for q in range(nd)])
        else:
            sig = ", ".join(f"int d{q}" for q in range(nd))
        L.append(f"  __device__ __forceinline__ const {b.typ} &operator()({sig}) const {{")
        expr = f"v[{len(sl) - 1}]"
        for k in range(len(sl) - 2, -1, -1):
            c, dx, dy, dz = sl[k]
            conds = [f"d0 == {dx}"] + ([f"d1 == {dy}"] if nd > 1 else []) + ([f"d2 == {dz}"] if nd > 2 else [])
            if b.multi:
                conds = [f"c == {max(c, 0)}"] + conds
            expr = f"({' && '.join(conds)}) ? v[{k}] : {expr}"
        L.append(f"    return {expr};")
        L.append("  }")
        # legacy (OPS_ACCn) kernels index with offset objects
        args_off = ", ".join(["o.x", "o.y", "o.z"][:nd])
        if b.multi:
            L.append(f"  __device__ __forceinline__ const {b.typ} &operator[](::opsmi_gen::OffMD o) const "
                     f"{{ return (*this)(o.d, {args_off}); }}")
        else:
            L.append(f"  __device__ __forceinline__ const {b.typ} &operator[](::opsmi_gen::Off o) const "
                     f"{{ return (*this)({args_off}); }}")
        L.append("};")
    L.append(f"__global__ __launch_bounds__(256) void opsg_fusenb_{F}(const opsa_{K1} A, const opsa_{K2} B"
             f"{''.join(', ' + p for p in ptrs)}) {{")
    L += red_init_lines(l1, "f") + red_init_lines(l2, "s")
    L.append("  ::opsmi_gen::for_each_point(A.L, [&](int x, int y, int z) {")
    L.append("    // the first loop at the point")
    L += point_lines(l1, "A", "f")
    # distinct offsets over the linked arguments' slots
    offs = sorted({q[1:] for i in slots for q in slots[i]})
    for i in slots:
        L.append(f"    opssh_{F}_{i} sh{i};")
    for o in offs:
        dx, dy, dz = o
        L.append(f"    {{  // offset ({dx}, {dy}, {dz})")
        fills = [(i, k, lk[i], q[0]) for i in slots for k, q in enumerate(slots[i]) if q[1:] == o]

        def val(src_acc, j, c):
            a = l1.args[j]
            idx = [str(dx)] + ([str(dy)] if nd > 1 else []) + ([str(dz)] if nd > 2 else [])
            if src_acc.startswith("fa"):
                if a.multi:
                    return f"{src_acc}({max(c, 0)}, {', '.join(idx)})"
                return f"{src_acc}({', '.join(idx)})"
            return f"{src_acc}.v[{max(c, 0)}]"
        if o == (0, 0, 0):
            for i, k, j, c in fills:
                L.append(f"      sh{i}.v[{k}] = {val(f'fa{j}', j, c)};")
        else:
            conds = [f"x + ({dx}) >= A.L.x0 && x + ({dx}) < A.L.x1"]
            if nd > 1:
                conds.append(f"y + ({dy}) >= A.L.y0 && y + ({dy}) < A.L.y1")
            if nd > 2:
                conds.append(f"z + ({dz}) >= A.L.z0 && z + ({dz}) < A.L.z1")
            L.append(f"      if ({' && '.join(conds)}) {{")
            L.append(f"        const int xo = x + ({dx}), yo = y + ({dy}), zo = z + ({dz});")
            call = []
            for j, a in enumerate(l1.args):
                if a.kind == "dat":
                    if a.acc == "OPS_READ":
                        L.append(f"        ::opsmi_gen::Acc<{a.typ}, {nd}, {'true' if a.multi else 'false'}, false> "
                                 f"ra{j}(fp{j}, A.d{j}, xo, yo, zo);")
                        call.append(f"ra{j}")
                    else:
                        L.append(f"        ::opsmi_gen::Cap<{a.typ}, {nd}, {a.dim if a.multi else 1}> cp{j}{{}};")
                        call.append(f"cp{j}")
                elif a.kind == "idx":
                    L.append(f"        int ridx{j}[3] = {{xo, yo, zo}};")
                    call.append(f"ridx{j}")
                elif a.kind == "gbl":
                    call.append(f"A.b.g{j}")
                else:
                    L.append(f"        {a.typ} rred{j}[{a.dim}];")
                    L.append(f"        for (int q = 0; q < {a.dim}; q++) rred{j}[q] = ::opsmi_gen::ident<{a.typ}, {a.acc}>();")
                    call.append(f"rred{j}")
            L.append(f"        A.b.k({', '.join(call)});")
            for i, k, j, c in fills:
                L.append(f"        sh{i}.v[{k}] = {val(f'cp{j}', j, c)};")
            L.append("      } else {")
            for i, k, j, c in fills:
                L.append(f"        sh{i}.v[{k}] = {val(f'fa{j}', j, c)};")
            L.append("      }")
        L.append("    }")
    L.append("    // the second loop, its linked arguments read from the held values")
    call = []
    for i, a in enumerate(l2.args):
        if a.kind == "dat":
            if i in lk:
                call.append(f"sh{i}")
            else:
                L.append(f"    ::opsmi_gen::Acc<{a.typ}, {nd}, {'true' if a.multi else 'false'}, false> "
                         f"sa{i}(sp{i}, B.d{i}, x, y, z);")
                call.append(f"sa{i}")
        elif a.kind == "idx":
            L.append(f"    int sidx{i}[3] = {{x, y, z}};")
            call.append(f"sidx{i}")
        elif a.kind == "gbl":
            call.append(f"B.b.g{i}")
        else:
            call.append(f"sred{i}")
    L.append(f"    B.b.k({', '.join(call)});")
    L.append("  });")
    L += red_store_lines(l1, "A", "f") + red_store_lines(l2, "B", "s")
    L.append("}")
    L.append(f"void opsfnb_{F}(const opsmi_gen_launch *L1, const void *blob1, const opsmi_gen_launch *L2, "
             "const void *blob2, void *stream) {")
    L.append(f"  opsa_{K1} A;")
    L.append(f"  opsfill_{K1}(A, L1, blob1);")
    L.append(f"  opsa_{K2} B;")
    L.append(f"  opsfill_{K2}(B, L2, blob2);")
    L.append(f"  hipLaunchKernelGGL(opsg_fusenb_{F}, dim3(L1->grid[0], L1->grid[1], L1->grid[2]), "
             f"dim3(::opsmi_gen::kWave, ::opsmi_gen::kRows), 0, (hipStream_t)stream, A, B"
             f"{launch_ptrs(l1, 'L1')}{launch_ptrs(l2, 'L2')});")
    L.append("}")
    flat = ", ".join(f"{i}, {j}" for i, j in links)
    reg = (f"{{ static const int lk[{2 * len(links)}] = {{{flat}}}; "
           f"ops_mi_gen_fused_nb(&opsl_{K1}, &opsl_{K2}, &opsfnb_{F}, lk, {len(links)}); }}")
    return "\n".join(L), reg


def adjacent_pairs(loops):
    """(first, second) loops of each pair of ops_par_loop calls that follow
    each other in the source, once per kernel pair, same dimension only."""
    seen, out = set(), []
    for a, b in zip(loops, loops[1:]):
        if a.ndim == b.ndim and (a.kernel, b.kernel) not in seen:
            seen.add((a.kernel, b.kernel))
            out.append((a, b))
    return out


def point_chains(loops, kpoints, n):
    """Runs of n ops_par_loop calls that follow each other in the source
    (same dimension, once per kernel sequence) in which every dat
    expression any call writes is accessed by all n kernels only at the
    point itself: the chains the library may run as one pass (it checks the
    dats themselves at run time)."""
    seen, out = set(), []
    for w in range(len(loops) - n + 1):
        ls = loops[w:w + n]
        key = tuple(l.kernel for l in ls)
        if key in seen or any(l.ndim != ls[0].ndim for l in ls):
            continue
        written = {a.expr for l in ls for a in l.args if a.kind == "dat" and a.acc != "OPS_READ"}
        ok = True
        for l in ls:
            for i, a in enumerate(l.args):
                if a.kind == "dat" and a.expr in written:
                    pts = kpoints[l.kernel][i]
                    if pts is None or any(tuple(q[1:]) != (0, 0, 0) for q in pts):
                        ok = False
        if ok:
            seen.add(key)
            out.append(ls)
    return out


def stub_code(loop, used, consts, points):
    K = loop.kernel
    n = len(loop.args)
    params = ", ".join(f"ops_arg arg{i}" for i in range(n))
    sig = f"void ops_par_loop_{K}(const char *name, ops_block block, int dim, int *range{', ' if n else ''}{params})"
    L = [f"extern \"C\" OPSMI_GEN_STUB {sig} {{"]
    if n:
        L.append(f"  const ops_arg args[{n}] = {{{', '.join(f'arg{i}' for i in range(n))}}};")
        types = ", ".join("OPSMI_GEN_NONE" if a.kind == "idx" else TYPES[a.typ][0] for a in loop.args)
        L.append(f"  static const int types[{n}] = {{{types}}};")
        counts = [a.dim if a.kind == "gbl" else (-1 if p is None else len(p)) for p, a in zip(points, loop.args)]
        flat = [v for p in points if p for pt in p for v in pt]
        L.append(f"  static const int npts[{n}] = {{{', '.join(map(str, counts))}}};")
        L.append(f"  static const int pts[{max(1, len(flat))}] = {{{', '.join(map(str, flat)) or '0'}}};")
    else:
        L.append("  const ops_arg *args = nullptr;")
        L.append("  const int *types = nullptr, *npts = nullptr, *pts = nullptr;")
    L.append(f"  opsb_{K} b;")
    L.append("  memset(&b, 0, sizeof(b));")
    for c in used:
        d, t = consts[c]
        L.append("  " + const_fetch(c, d, t))
    for i, a in enumerate(loop.args):
        if a.kind == "gbl":
            L.append(f"  if (arg{i}.data && arg{i}.elem_size == {TYPES[a.typ][1]} && arg{i}.dim > 0)")
            L.append(f"    memcpy(b.g{i}, arg{i}.data, (size_t)(arg{i}.dim < {a.dim} ? arg{i}.dim : {a.dim}) * "
                     f"{TYPES[a.typ][1]});")
    L.append(f"  ops_mi_gen_par_loop(name, block, dim, range, args, {n}, types, npts, pts, &opsl_{K}, &b, sizeof(b));")
    L.append("}")
    return "\n".join(L), sig


def translate(app, out_dir, inc_dirs=(), program=None):
    """Translates one source file; program: what the other files of the same
    program provide (Program), or None for a file on its own."""
    inc_dirs = list(inc_dirs)
    src = open(app).read()
    m = mask(src)
    base = os.path.splitext(os.path.basename(app))[0]
    loops = find_loops(src, m, app)
    incs = local_includes(app, src, m, inc_dirs)
    files = [(app, src, m)]
    for _, _, p, _ in incs:
        s = open(p).read()
        files.append((p, s, mask(s)))
    consts = dict(program.consts) if program else {}
    for _, s, mm in files:
        consts.update(find_consts(s, mm))
    # kernels: this file and its headers first, then the rest of the program
    kfiles = files + ([f for f in program.files if f[0] not in {q[0] for q in files}] if program else [])
    kernels, order = {}, []
    for lp in loops:
        if lp.kernel in kernels:
            prev = kernels[lp.kernel][0]
            if [(a.kind, a.typ, a.acc, a.dim, a.multi) for a in prev.args] != \
               [(a.kind, a.typ, a.acc, a.dim, a.multi) for a in lp.args] or prev.ndim != lp.ndim:
                raise TranslateError(f"kernel {lp.kernel} is called with different argument lists")
            continue
        kernels[lp.kernel] = (lp, find_kernel(lp.kernel, kfiles))
        order.append(lp.kernel)

    # kernels file: the application's other local headers (user types and
    # helper functions) come first, with OPS_FUN_PREFIX making the helpers
    # device functions, as the reference's generated code does
    kernel_files = {kern.path for _, kern in kernels.values()} | (program.kernel_headers if program else set())
    ops_names = {"ops_seq_v2.h", "ops_seq.h", "ops_lib_core.h"}
    # headers of kernels no loop calls (they take ACC<T> parameters) stay in
    # the application only: their host functions would be defined twice
    acc_headers = {p for _, _, p, _ in incs if re.search(r"\bACC\s*<", mask(open(p).read()))}
    user_incs = [name for _, _, p, name in incs
                 if p not in kernel_files and p not in acc_headers and os.path.basename(name) not in ops_names
                 and not defines_globals(open(p).read()) and not defines_functions(open(p).read())]
    H = [f"// Generated by ops_amd/translator/opsmi_translate.py from {os.path.basename(app)}; do not edit.",
         "#include <math.h>", "#include <string.h>", "", "#include \"ops_mi355x_device.h\"", "",
         "#ifndef OPS_FUN_PREFIX", "#define OPS_FUN_PREFIX __host__ __device__", "#endif"]
    H += [f"#include \"{name}\"" for name in user_incs]
    # stubs are weak: an application whose source files each call the same
    # kernel links one of the (identical) stubs
    H += ["", "#define OPSMI_GEN_STUB __attribute__((visibility(\"hidden\"), weak))", "", "namespace {"]
    stubs, sigs, kpoints = [], [], {}
    for k in order:
        lp, kern = kernels[k]
        code, used, points = kernel_code(lp, kern, consts)
        kpoints[k] = points
        H.append(code)
        s, sig = stub_code(lp, used, consts, points)
        stubs.append(s)
        sigs.append(sig)
    regs = []
    # adjacent pairs: the same one-pass form as the runs below
    for a, b in adjacent_pairs(loops):
        got = chain_code([kernels[a.kernel][0], kernels[b.kernel][0]], [a, b])
        if got:
            H.append(got[0])
            regs.append(got[1])
    # runs of three and four calls touching what they write only at the point
    for n in (4, 3):
        for ls in point_chains(loops, kpoints, n):
            got = chain_code([kernels[l.kernel][0] for l in ls], ls)
            if got:
                H.append(got[0])
                regs.append(got[1])
    # pairs whose second loop reads the first's output at neighbours:
    # adjacent calls and the last and first call of a loop body, matched by
    # the dat expressions as written (the library checks the dats at run time)
    seen = set()
    for a, b in adjacent_pairs(loops) + loop_body_pairs(src, m, loops):
        if a.ndim != b.ndim or (a.kernel, b.kernel) in seen or a.kernel == b.kernel:
            continue
        seen.add((a.kernel, b.kernel))
        l1, l2 = kernels[a.kernel][0], kernels[b.kernel][0]
        # match expressions on these calls, the argument lists on the kernels' first calls
        links = nb_links(a, b, kpoints[b.kernel])
        if not links or any(p is None for p, x in zip(kpoints[a.kernel], l1.args) if x.kind == "dat"
                            and x.acc != "OPS_READ"):
            continue
        code, reg = nb_fused_code(l1, l2, kpoints[a.kernel], kpoints[b.kernel], links)
        H.append(code)
        regs.append(reg)
    if regs:
        H.append("struct opsr_fused_pairs {")
        H.append("  opsr_fused_pairs() {")
        H += [f"    {r}" for r in regs]
        H.append("  }")
        H.append("} opsr_fused_pairs_;")
    H.append("}  // namespace")
    H.append("")
    H.extend(stubs)
    os.makedirs(out_dir, exist_ok=True)
    with open(os.path.join(out_dir, base + "_opsmi_kernels.hip"), "w") as f:
        f.write("\n".join(H) + "\n")
    guard = re.sub(r"\W", "_", base).upper() + "_OPSMI_KERNELS_H"
    D = [f"// Generated by ops_amd/translator/opsmi_translate.py from {os.path.basename(app)}; do not edit.",
         f"#ifndef {guard}", f"#define {guard}", "#include \"ops_mi355x.h\"", "#ifdef __cplusplus", "extern \"C\" {",
         "#endif"]
    for sig in sigs:
        D.append(f"__attribute__((visibility(\"hidden\"))) {sig};")
    D += ["#ifdef __cplusplus", "}", "#endif", "#endif"]
    with open(os.path.join(out_dir, base + "_opsmi_kernels.h"), "w") as f:
        f.write("\n".join(D) + "\n")

    # rewritten application: edits from the back so offsets stay valid
    edits = []
    for lp in loops:
        edits.append((lp.span[0], lp.span[1], f"ops_par_loop_{lp.kernel}("))
    for k in order:
        kern = kernels[k][1]
        if kern.path == app:
            edits.append((kern.span[0], kern.span[1], f"/* kernel {k}: generated into {base}_opsmi_kernels.hip */"))
    for a, b, p, name in incs:
        if p in kernel_files:
            edits.append((a, b, f"/* #include \"{name}\": kernels generated into {base}_opsmi_kernels.hip */"))
    ops_inc = re.compile(r'#\s*include\s*[<"](ops_seq_v2\.h|ops_seq\.h|ops_lib_core\.h)[>"]')
    first_ops = None
    for mo in ops_inc.finditer(src):
        if m[mo.start()] != "#":
            continue
        edits.append((mo.start(), mo.end(), "#include \"ops_mi355x.h\""))
        first_ops = mo.end() if first_ops is None else first_ops
    out = src
    for a, b, t in sorted(edits, key=lambda e: -e[0]):
        out = out[:a] + t + out[b:]
    decl = f"#include \"{base}_opsmi_kernels.h\"\n"
    if first_ops is not None:
        # after the (rewritten) OPS include: find it again in the output
        mo = re.search(r'#include "ops_mi355x\.h"\n?', out)
        out = out[:mo.end()] + ("" if out[mo.end() - 1] == "\n" else "\n") + decl + out[mo.end():]
    else:
        out = "#include \"ops_mi355x.h\"\n" + decl + out
    with open(os.path.join(out_dir, base + "_ops.cpp"), "w") as f:
        f.write(out)
    return order


ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build_shared(apps, out_so, gen_dir, inc_dirs=(), hipcc="/opt/rocm/bin/hipcc", jobs=8, executable=False):
    """Translates apps into gen_dir and links them with their generated
    kernels into the shared object out_so (or, with executable=True, the
    program out_so), linked against ops_amd/libopsmi.so with an rpath to
    it."""
    import subprocess
    from concurrent.futures import ThreadPoolExecutor
    inc = [os.path.join(ROOT, "include"), gen_dir] + [os.path.dirname(os.path.abspath(a)) for a in apps] + list(inc_dirs)
    iflags = [f"-I{d}" for d in inc]
    cmds, objs = [], []
    # a program's files share its constants; separate applications linked
    # into one library keep their own
    prog = Program(apps, inc_dirs) if executable else None
    for app in apps:
        translate(app, gen_dir, inc_dirs, prog)
        base = os.path.splitext(os.path.basename(app))[0]
        for src, obj, lang in ((f"{base}_opsmi_kernels.hip", f"{base}_k.o", ["-x", "hip"]),
                               (f"{base}_ops.cpp", f"{base}_a.o", [])):
            o = os.path.join(gen_dir, obj)
            cmds.append([hipcc] + HIP_FLAGS + iflags + lang + ["-c", os.path.join(gen_dir, src), "-o", o])
            objs.append(o)
    with ThreadPoolExecutor(max(1, jobs)) as ex:
        for r in ex.map(lambda c: subprocess.run(c, capture_output=True, text=True), cmds):
            if r.returncode != 0:
                raise TranslateError(f"compile failed: {' '.join(r.args)}\n{r.stderr}")
    libdir = os.path.join(ROOT, "ops_amd")
    rpath = os.path.relpath(libdir, os.path.dirname(os.path.abspath(out_so)))
    link = [hipcc] + HIP_FLAGS + ([] if executable else ["-shared"]) + objs + [f"-L{libdir}", "-lopsmi",
                                                       f"-Wl,-rpath,$ORIGIN/{rpath}", "-o", out_so]
    r = subprocess.run(link, capture_output=True, text=True)
    if r.returncode != 0:
        raise TranslateError(f"link failed: {' '.join(link)}\n{r.stderr}")
    return out_so


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("apps", nargs="+")
    ap.add_argument("-o", "--out", default=".")
    ap.add_argument("-I", dest="inc", action="append", default=[])
    ap.add_argument("--shared", metavar="LIB.so", help="also compile and link everything into LIB.so")
    ap.add_argument("--exe", metavar="PROGRAM", help="also compile and link everything into PROGRAM")
    a = ap.parse_args(argv)
    try:
        if a.shared or a.exe:
            out = a.exe or a.shared
            build_shared(a.apps, out, a.out, a.inc, executable=bool(a.exe))
            print(f"{len(a.apps)} applications -> {out}")
        else:
            for app in a.apps:
                ks = translate(app, a.out, a.inc)
                print(f"{app}: {len(ks)} kernels -> {a.out}")
    except TranslateError as e:
        print(f"opsmi_translate: error: {e}", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
