"""MI355X-native per-step gravity path with the module layout of Cosmo-NBody's
``src`` package (src/kernels, src/core, src/utils), so that callers written
against the reference (``from src.kernels.morton import compute_codes``) run
unchanged once this directory's parent is on ``sys.path``."""
