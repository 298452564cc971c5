"""Every name the Python sources read is bound somewhere in scope: a local,
an enclosing function's, a module-level one or a builtin.  A misspelt or
moved name otherwise only fails when its line runs, on a GPU box and, for
bench.py, at the end of a long run.  The names spimage.py binds at import
time through _bind("name", ...) (or its alias=) count as module-level."""
import ast
import builtins
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SOURCES = ([ROOT / "bench.py", ROOT / "__graft_entry__.py"] + sorted((ROOT / "libspimage_amd").glob("*.py"))
           + sorted((ROOT / "oracle").glob("*.py")) + sorted((ROOT / "tests").glob("*.py"))
           + sorted((ROOT / "tests" / "golden").glob("*.py")))

SCOPES = (ast.FunctionDef, ast.AsyncFunctionDef, ast.Lambda, ast.ClassDef, ast.ListComp, ast.SetComp,
          ast.DictComp, ast.GeneratorExp)


def _bound_here(node):
    """Names a scope binds: its parameters and every name stored, imported,
    defined or declared global/nonlocal directly in it (not in nested
    scopes, except the names nested defs and classes bind for it)."""
    names = set()
    if isinstance(node, (ast.FunctionDef, ast.AsyncFunctionDef, ast.Lambda)):
        a = node.args
        for arg in a.posonlyargs + a.args + a.kwonlyargs:
            names.add(arg.arg)
        if a.vararg:
            names.add(a.vararg.arg)
        if a.kwarg:
            names.add(a.kwarg.arg)
    if isinstance(node, (ast.ListComp, ast.SetComp, ast.DictComp, ast.GeneratorExp)):
        for gen in node.generators:
            for n in ast.walk(gen.target):
                if isinstance(n, ast.Name):
                    names.add(n.id)
        return names
    body = node.body if isinstance(node.body, list) else [node.body]
    stack = list(body)
    while stack:
        n = stack.pop()
        if isinstance(n, (ast.FunctionDef, ast.AsyncFunctionDef, ast.ClassDef)):
            names.add(n.name)
            stack.extend(n.decorator_list)
            continue
        if isinstance(n, SCOPES):
            continue
        if isinstance(n, ast.Name) and isinstance(n.ctx, (ast.Store, ast.Del)):
            names.add(n.id)
        elif isinstance(n, (ast.Import, ast.ImportFrom)):
            for al in n.names:
                names.add((al.asname or al.name).split(".")[0])
        elif isinstance(n, (ast.Global, ast.Nonlocal)):
            names.update(n.names)
        elif isinstance(n, ast.ExceptHandler) and n.name:
            names.add(n.name)
        elif isinstance(n, ast.MatchAs) and n.name:
            names.add(n.name)
        elif (isinstance(n, ast.Call) and isinstance(n.func, ast.Name) and n.func.id == "_bind" and n.args
              and isinstance(n.args[0], ast.Constant) and isinstance(n.args[0].value, str)):
            # _bind(name, ..., alias=other) binds `other` instead
            alias = [k.value.value for k in n.keywords if k.arg == "alias" and isinstance(k.value, ast.Constant)]
            names.add(alias[0] if alias else n.args[0].value)
        stack.extend(ast.iter_child_nodes(n))
    return names


def _unbound(tree):
    missing = []
    module_names = _bound_here(tree) | set(dir(builtins)) | {"__file__", "__name__", "__doc__"}

    def visit(node, chain):
        for child in ast.iter_child_nodes(node):
            if isinstance(child, SCOPES):
                inner = chain + [_bound_here(child)]
                # defaults and decorators are read in the enclosing scope
                if isinstance(child, (ast.FunctionDef, ast.AsyncFunctionDef, ast.Lambda)):
                    for d in child.args.defaults + [d for d in child.args.kw_defaults if d is not None]:
                        check(d, chain)
                if isinstance(child, (ast.ListComp, ast.SetComp, ast.DictComp, ast.GeneratorExp)):
                    # the first iterable is evaluated in the enclosing scope
                    check(child.generators[0].iter, chain)
                visit(child, inner)
            else:
                if isinstance(child, ast.Name) and isinstance(child.ctx, ast.Load):
                    if not any(child.id in s for s in chain) and child.id not in module_names:
                        missing.append((child.id, child.lineno))
                visit(child, chain)

    def check(expr, chain):
        visit(ast.Expression(body=expr), chain)

    visit(tree, [])
    return missing


@pytest.mark.parametrize("path", SOURCES, ids=lambda p: str(p.relative_to(ROOT)))
def test_every_name_read_is_bound(path):
    tree = ast.parse(path.read_text(), filename=str(path))
    assert _unbound(tree) == []


def test_the_check_finds_an_unbound_name():
    src = ("import os\n"
           "def f(a, *, k=os.sep):\n"
           "    b = [x for x in range(a)]\n"
           "    return {'n': len(shapes), 'b': b, 'k': k}\n")
    assert _unbound(ast.parse(src)) == [("shapes", 4)]
