"""Package version; the API follows pynear 2.2.0 (pynear/_version.py:3-8)."""

__version__ = "0.1.0"
