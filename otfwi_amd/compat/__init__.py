"""Drop-in replacements for the reference's top-level helper modules
(`geometry`, `utils`) without scikit-image or plotting dependencies.  Put
this directory on sys.path (or PYTHONPATH) so that `from geometry import *`
in the reference's setups/*.py resolves here."""
