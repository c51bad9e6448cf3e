"""MI355X-native covariant-tensor partial-wave likelihood (drop-in for ctpwa's analysis path).

    import ctpwa_amd as ctpwa
    ana = ctpwa.analysis("config.yml")
    nll = ana.getNLL(params); grad, = torch.autograd.grad(nll, params)
"""
from .config import ConfigParser, ConfigError  # noqa: F401
from .model import Plan, build_plan  # noqa: F401

__version__ = "0.1.0"


def __getattr__(name):
    # the analysis class pulls in torch and the native library lazily
    if name in ("analysis", "NLLFunction"):
        import importlib
        return getattr(importlib.import_module("._analysis", __name__), name)
    raise AttributeError(name)
