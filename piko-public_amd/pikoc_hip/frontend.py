"""Source frontend of pikoc_hip: reads a Piko pipe header and the stage
headers it includes and builds the pipe summary.

pikoc builds the same summary with a Clang AST pass (pikoc/src/Frontend/
StageASTVisitor.cpp:100-447, PipeASTVisitor); the Clang 3.2 libraries it
needs are not available, so this frontend works on the source text with a
small preprocessor (#define / #ifdef / #ifndef / #if defined / #else /
#endif, quoted #include) and pattern matching of the forms a Piko pipe uses:

  class S : public Stage<binX, binY, threadsPerTile, InPrim, OutPrim> { ... }
      assignBin(...) { assignToBin(p, PREVIOUS_BINS | BOUNDING_BOX | POSITION | <expr>); }
      schedule(...)  { specifySchedule(LOAD_BALANCE | ROUND_ROBIN | SERIAL | ALL[, tileSplitSize]);
                       specifyWait(BATCH | END_STAGE | END_BIN); }
      process(...)   { specifyMaxOutPrims(n); ... }
  class P : public PikoPipe { S1 a; S2 b; ... P() { pikoConnect(a, b, outPort, inPort); }
      void run(ConstantState&, MutableState&, Input*, int); ... }
  struct T : public Primitive { field declarations }
"""
import ast
import operator
import os
import re

from .summary import PipeSummary, StageSummary

GENERATED = ("__pikoDefines.h", "__pikoCompiledPipe.h")


class FrontendError(Exception):
    pass


# ---------------------------------------------------------------------------
# preprocessor

_DIRECTIVE = re.compile(r"^\s*#\s*(\w+)\s*(.*)$")


def _strip_comments(text):
    text = re.sub(r"/\*.*?\*/", lambda m: "\n" * m.group(0).count("\n"), text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def preprocess(path, include_dirs=(), defines=None, _seen=None):
    """Returns (text, macros): the source with quoted includes inlined and
    inactive conditional blocks removed; macros maps names to their values."""
    macros = dict(defines or {})
    seen = _seen if _seen is not None else set()
    out = []

    def resolve(name, here):
        for d in (os.path.dirname(here),) + tuple(include_dirs):
            p = os.path.join(d, name)
            if os.path.exists(p):
                return os.path.abspath(p)
        return None

    def active(stack):
        return all(s[0] for s in stack)

    def run(p):
        if p in seen:
            return
        seen.add(p)
        with open(p) as f:
            lines = _strip_comments(f.read()).split("\n")
        stack = []  # (taking, any_taken)
        k = 0
        while k < len(lines):
            line = lines[k]
            while line.endswith("\\") and k + 1 < len(lines):
                k += 1
                line = line[:-1] + " " + lines[k]
            k += 1
            m = _DIRECTIVE.match(line)
            if not m:
                if active(stack):
                    out.append(line)
                continue
            d, rest = m.group(1), m.group(2).strip()
            if d in ("ifdef", "ifndef"):
                cond = (rest.split()[0] in macros) == (d == "ifdef") if rest else False
                stack.append([cond, cond])
            elif d == "if":
                cond = _eval_if(rest, macros)
                stack.append([cond, cond])
            elif d == "elif":
                if not stack:
                    raise FrontendError("#elif without #if in %s" % p)
                cond = (not stack[-1][1]) and _eval_if(rest, macros)
                stack[-1][0] = cond
                stack[-1][1] = stack[-1][1] or cond
            elif d == "else":
                if not stack:
                    raise FrontendError("#else without #if in %s" % p)
                stack[-1][0] = not stack[-1][1]
                stack[-1][1] = True
            elif d == "endif":
                if not stack:
                    raise FrontendError("#endif without #if in %s" % p)
                stack.pop()
            elif not active(stack):
                continue
            elif d == "define":
                mm = re.match(r"(\w+)(\([^)]*\))?\s*(.*)$", rest)
                if mm and not mm.group(2):
                    macros[mm.group(1)] = mm.group(3).strip()
            elif d == "undef":
                macros.pop(rest.split()[0], None)
            elif d == "include":
                mi = re.match(r'"([^"]+)"', rest)
                if mi and os.path.basename(mi.group(1)) not in GENERATED:
                    q = resolve(mi.group(1), p)
                    if q is not None and not mi.group(1).startswith("piko/"):
                        run(q)
        if stack:
            raise FrontendError("unterminated conditional in %s" % p)

    run(os.path.abspath(path))
    return "\n".join(out), macros


def _eval_if(expr, macros):
    expr = re.sub(r"defined\s*\(\s*(\w+)\s*\)", lambda m: "1" if m.group(1) in macros else "0", expr)
    expr = re.sub(r"defined\s+(\w+)", lambda m: "1" if m.group(1) in macros else "0", expr)
    try:
        return bool(eval_int(expr, macros))
    except FrontendError:
        return False


_OPS = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul, ast.FloorDiv: operator.floordiv,
        ast.Div: operator.floordiv, ast.LShift: operator.lshift, ast.RShift: operator.rshift,
        ast.BitOr: operator.or_, ast.BitAnd: operator.and_, ast.Mod: operator.mod}


def expand(expr, macros, depth=0):
    if depth > 32:
        raise FrontendError("macro expansion too deep in %r" % expr)
    def sub(m):
        name = m.group(0)
        return "(" + expand(macros[name], macros, depth + 1) + ")" if name in macros else name
    return re.sub(r"[A-Za-z_]\w*", sub, expr)


def eval_number(expr, macros):
    """Value of an integer or floating constant expression after macro
    expansion (+ - * / % << >> | & and parentheses only)."""
    e = expand(expr, macros)
    e = re.sub(r"(\d+\.\d*|\.\d+|\d+)[fFuUlL]+\b", r"\1", e)
    try:
        tree = ast.parse(e.strip(), mode="eval")
    except SyntaxError as exc:
        raise FrontendError("not a constant expression: %r" % expr) from exc

    def ev(n):
        if isinstance(n, ast.Expression):
            return ev(n.body)
        if isinstance(n, ast.Constant) and isinstance(n.value, (int, float)):
            return n.value
        if isinstance(n, ast.UnaryOp) and isinstance(n.op, (ast.USub, ast.UAdd)):
            v = ev(n.operand)
            return -v if isinstance(n.op, ast.USub) else v
        if isinstance(n, ast.BinOp) and type(n.op) in _OPS:
            a, b = ev(n.left), ev(n.right)
            if isinstance(n.op, ast.Div) and (isinstance(a, float) or isinstance(b, float)):
                return a / b
            return _OPS[type(n.op)](a, b)
        raise FrontendError("not a constant expression: %r" % expr)
    return ev(tree)


def eval_int(expr, macros):
    v = eval_number(expr, macros)
    if int(v) != v:
        raise FrontendError("not an integer: %r" % expr)
    return int(v)


# ---------------------------------------------------------------------------
# structure

def _block(text, open_pos):
    """Index just past the brace block starting at text[open_pos] == '{'."""
    depth = 0
    for k in range(open_pos, len(text)):
        if text[k] == "{":
            depth += 1
        elif text[k] == "}":
            depth -= 1
            if depth == 0:
                return k + 1
    raise FrontendError("unbalanced braces")


def _flat(text):
    """text with every brace block (function bodies, nested types) removed."""
    out, k = [], 0
    while k < len(text):
        if text[k] == "{":
            k = _block(text, k)
            out.append(";")
        else:
            out.append(text[k])
            k += 1
    return "".join(out)


def _split_args(s):
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "<([":
            depth += 1
        elif ch in ">)]":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur.strip())
    return out


def _method_body(body, name):
    m = re.search(r"\b%s\s*\([^)]*\)\s*(const\s*)?\{" % name, body)
    if not m:
        return None
    start = m.end() - 1
    return body[start + 1:_block(body, start) - 1]


def _calls(body, fn):
    return [_split_args(m.group(1)) for m in re.finditer(r"\b%s\s*\(([^;]*)\)\s*;" % fn, body)]


def _statements(body):
    return [s.strip() for s in body.split(";") if s.strip()]


def normalise(code):
    """Source text with layout removed: whitespace runs become one space, and
    spaces next to punctuation go, so re-indenting or re-wrapping a body does
    not change it, while changing, adding or removing an identifier, number
    or operator does (comments are already gone)."""
    code = re.sub(r"\s+", " ", code).strip()
    return re.sub(r" ?([^\w ]) ?", r"\1", code)


def parse_stage(name, args, body, macros):
    if len(args) != 5:
        raise FrontendError("Stage<> of %s needs 5 template arguments" % name)
    s = StageSummary(name=name, type=name)
    # the class body and each phase's body as the backend sees them: after
    # the preprocessor (inactive branches removed), macros not expanded
    s.source = {"class": normalise(body)}
    for meth in ("assignBin", "schedule", "process"):
        mb = _method_body(body, meth)
        s.source[meth] = None if mb is None else normalise(mb)
    s.binsize = (eval_int(args[0], macros), eval_int(args[1], macros))
    s.threads_per_tile = eval_int(args[2], macros)
    s.prim_in, s.prim_out = args[3], args[4]
    ab = _method_body(body, "assignBin")
    if ab is None:
        s.assign_policy = "Empty"
    else:
        calls = _calls(ab, r"(?:this->)?assignToBin")
        pols = [c[1] for c in calls if len(c) == 2 and c[1] in ("PREVIOUS_BINS", "BOUNDING_BOX", "POSITION")]
        manual = [c for c in calls if not (len(c) == 2 and c[1] in ("PREVIOUS_BINS", "BOUNDING_BOX", "POSITION"))]
        if pols and manual:
            raise FrontendError("%s: assignBin policy and manual binning together" % name)
        if len(pols) > 1:
            raise FrontendError("%s: more than one assignBin policy" % name)
        if pols:
            s.assign_policy = {"PREVIOUS_BINS": "InBin", "BOUNDING_BOX": "BoundingBox", "POSITION": "Position"}[pols[0]]
            s.assign_trivial = len(_statements(ab)) == 1
        elif manual:
            s.assign_policy = "Custom"
        else:
            s.assign_policy = "Empty"
    sb = _method_body(body, "schedule")
    if sb is not None:
        scheds = _calls(sb, "specifySchedule")
        waits = _calls(sb, "specifyWait")
        if len(scheds) > 1 or len(waits) > 1:
            raise FrontendError("%s: more than one schedule or wait policy" % name)
        if scheds:
            pol = scheds[0][0]
            table = {"LOAD_BALANCE": "LoadBalance", "ROUND_ROBIN": "DirectMap", "SERIAL": "Serialize", "ALL": "All"}
            if pol not in table:
                raise FrontendError("%s: schedule policy %s not recognised" % (name, pol))
            s.sched_policy = table[pol]
            if len(scheds[0]) > 1 and pol in ("LOAD_BALANCE", "ALL"):
                s.tile_split_size = eval_int(scheds[0][1], macros)
        else:
            s.sched_policy = "Custom"
        if waits:
            table = {"BATCH": "Batch", "END_STAGE": "EndStage", "END_BIN": "EndBin"}
            if waits[0][0] not in table:
                raise FrontendError("%s: wait policy %s not recognised" % (name, waits[0][0]))
            s.wait_policy = table[waits[0][0]]
        n = len(_statements(sb))
        s.sched_trivial = bool(scheds) and n == len(scheds) + len(waits)
    pb = _method_body(body, "process")
    if pb is not None:
        mo = _calls(pb, "specifyMaxOutPrims")
        if mo:
            s.max_out_prims = eval_int(mo[0][0], macros)
        s.process_policy = "Custom" if _statements(pb) else "Empty"
    return s


def parse_struct_fields(text, name):
    m = re.search(r"\bstruct\s+%s\b[^{;]*\{" % re.escape(name), text)
    if not m:
        return None
    body = text[m.end():_block(text, m.end() - 1) - 1]
    fields = []
    for st in _statements(body):
        fm = re.match(r"^([\w:]+(?:\s*\*)?)\s+(\w+)\s*(?:\[\s*([^\]]+)\s*\])?$", st)
        if fm and "(" not in st:
            fields.append((fm.group(1).replace(" ", ""), fm.group(2), fm.group(3)))
    return fields


def parse_pipe(path, include_dirs=(), defines=None):
    # stage bodies sit in #ifdef __PIKOC_DEVICE__ blocks, as pikoc reads them
    defines = dict(defines or {})
    defines.setdefault("__PIKOC_DEVICE__", "1")
    text, macros = preprocess(path, include_dirs, defines)
    stages = {}
    for m in re.finditer(r"\bclass\s+(\w+)\s*:\s*public\s+Stage\s*<", text):
        lt = m.end()
        depth, k = 1, lt
        while depth:
            if text[k] == "<":
                depth += 1
            elif text[k] == ">":
                depth -= 1
            k += 1
        args = _split_args(text[lt:k - 1])
        ob = text.index("{", k)
        body = text[ob + 1:_block(text, ob) - 1]
        stages[m.group(1)] = parse_stage(m.group(1), args, body, macros)
    pm = list(re.finditer(r"\bclass\s+(\w+)\s*:\s*public\s+PikoPipe\s*\{", text))
    if len(pm) != 1:
        raise FrontendError("expected exactly one PikoPipe class, found %d" % len(pm))
    pm = pm[0]
    body = text[pm.end():_block(text, pm.end() - 1) - 1]
    ps = PipeSummary(name=pm.group(1), filename=os.path.basename(path))
    ps.macros = macros
    for st_text in _statements(_flat(body)):
        st_text = re.sub(r"^(public|private|protected)\s*:\s*", "", st_text)
        mm = re.match(r"^(\w+)\s+(\w+)$", st_text)
        if mm and mm.group(1) in stages:
            ps.add_stage(stages[mm.group(1)].copy_as(mm.group(2)))
    if not ps.stages:
        raise FrontendError("pipe %s declares no stages" % ps.name)
    ctor = re.search(r"\b%s\s*\(\s*\)\s*(?::[^{]*)?\{" % re.escape(ps.name), body)
    if not ctor:
        raise FrontendError("pipe %s has no default constructor with pikoConnect calls" % ps.name)
    cbody = body[ctor.end():_block(body, ctor.end() - 1) - 1]
    for args in _calls(cbody, "pikoConnect"):
        if len(args) != 4:
            raise FrontendError("pikoConnect takes 4 arguments")
        ps.connect(args[0], args[1], eval_int(args[2], macros), eval_int(args[3], macros))
    rm = re.search(r"\bvoid\s+run\s*\(\s*(\w+)\s*&\s*\w*\s*,\s*(\w+)\s*&\s*\w*\s*,\s*(\w+)\s*\*\s*\w*\s*,\s*int",
                   body)
    if not rm:
        raise FrontendError("pipe %s must declare run(ConstantState&, MutableState&, Input*, int)" % ps.name)
    ps.const_state_type, ps.mutable_state_type, ps.input_type = rm.groups()
    for meth in ("allocate", "prepare", "run_single", "destroy"):
        if not re.search(r"\bvoid\s+%s\s*\(" % meth, body):
            raise FrontendError("pipe %s must declare %s()" % (ps.name, meth))
    ps.input_fields = parse_struct_fields(text, ps.input_type)
    ps.process_links()
    return ps
