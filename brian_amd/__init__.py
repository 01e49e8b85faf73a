"""brian_amd: MI355X (gfx950) backend for Brian 1.4.4's clock-driven network step."""
__version__ = "0.1.0"
