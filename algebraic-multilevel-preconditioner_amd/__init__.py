"""MI355X apply path of the two-level overlapping Schwarz preconditioner (schwarz2lvl).

Import it through :func:`load` (the directory name is not a Python identifier)::

    import importlib.util, sys
    amp = load()              # registers the package as ``amp_amd``
"""
