"""Helpers around the operators: FDK filtering and subset ordering."""
from .filtering import filtering, ramp_filter  # noqa: F401
from .order_subsets import order_subsets  # noqa: F401
