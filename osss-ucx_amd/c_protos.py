"""Function prototypes of a C header, as data: name -> normalised
"return(param, param, ...)" with parameter names dropped.

gen_pshmem_h.py builds include/pshmem.h from include/shmem.h with it, and
the boundary tests compare include/shmem.h with the recorded prototypes of
the reference's include/shmem/api.h through the same code.  The header is
preprocessed with gcc -E; the declarations are then split at top-level
semicolons and braces."""
import re
import subprocess

_KEYWORDS = {
    "void", "char", "short", "int", "long", "float", "double", "signed",
    "unsigned", "const", "volatile", "restrict", "__restrict", "_Complex",
    "_Bool", "struct", "enum", "union", "inline", "static", "extern",
}


def preprocess(header, include_dirs, defines=()):
    cmd = ["gcc", "-E", "-P", "-std=gnu11"]
    for d in include_dirs:
        cmd += ["-I", d]
    for d in defines:
        cmd += ["-D" + d]
    cmd.append(header)
    return subprocess.run(cmd, check=True, capture_output=True,
                          text=True).stdout


def _strip_attributes(s):
    out, i = [], 0
    while True:
        j = s.find("__attribute__", i)
        if j < 0:
            out.append(s[i:])
            return "".join(out)
        out.append(s[i:j])
        k = s.index("(", j)
        depth = 0
        while True:
            if s[k] == "(":
                depth += 1
            elif s[k] == ")":
                depth -= 1
                if depth == 0:
                    break
            k += 1
        i = k + 1


def _tokens(s):
    return re.findall(r"[A-Za-z_]\w*|\.\.\.|\S", s)


def _param_type(p):
    t = _tokens(p)
    if len(t) >= 2 and re.match(r"[A-Za-z_]\w*$", t[-1]) and \
            t[-1] not in _KEYWORDS and t[-2] not in ("struct", "enum",
                                                      "union"):
        t = t[:-1]
    return " ".join(t)


def _split_params(s):
    parts, depth, cur = [], 0, []
    for ch in s:
        if ch == "," and depth == 0:
            parts.append("".join(cur))
            cur = []
            continue
        depth += ch == "("
        depth -= ch == ")"
        cur.append(ch)
    parts.append("".join(cur))
    return [p.strip() for p in parts if p.strip()]


PUBLIC_RE = r"^(p?shmem|sh|start_pes|_my_pe|_num_pes)"


def prototypes(text, prefix_re=PUBLIC_RE):
    """{name: 'ret(params)'} for every function declaration (not
    definition, typedef or variable) whose name matches prefix_re."""
    text = _strip_attributes(text)
    # drop function bodies (static inline helpers) and braces of
    # struct/enum definitions: keep only top-level statements
    stmts, depth, cur = [], 0, []
    for ch in text:
        if ch == "{":
            depth += 1
            continue
        if ch == "}":
            depth -= 1
            if depth == 0:
                cur = []   # a body or a type definition ends here
            continue
        if depth:
            continue
        if ch == ";":
            stmts.append("".join(cur))
            cur = []
        else:
            cur.append(ch)
    out = {}
    for st in stmts:
        st = " ".join(st.split())
        if not st or st.startswith("typedef"):
            continue
        m = re.match(r"^(?:extern\s+)?(.*?)\b([A-Za-z_]\w*)\s*\((.*)\)$", st)
        if m is None or "(" in m.group(1):
            continue
        ret, name, params = m.groups()
        if not re.match(prefix_re, name):
            continue
        ret = " ".join(_tokens(ret.replace("extern", "")))
        ps = [_param_type(p) for p in _split_params(params)]
        out[name] = "%s(%s)" % (ret, ", ".join(ps))
    return out
