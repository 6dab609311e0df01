"""Plugins that turn external layout descriptions into engine inputs."""
from .gdsfactory import component_to_eps

__all__ = ["component_to_eps"]
