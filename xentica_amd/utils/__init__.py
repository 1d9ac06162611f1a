"""Small utilities (``xentica/utils``)."""
