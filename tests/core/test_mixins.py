"""Mixins (``tests/core/test_mixins.py`` of the reference)."""
import pytest

from xentica_amd.core.mixins import BscaDetectorMixin
from xentica_amd.core.exceptions import XenticaException


def test_not_detected():
    class Dummy(BscaDetectorMixin):
        @property
        def test_bsca(self):
            return self.bsca
    with pytest.raises(XenticaException):
        Dummy().test_bsca
