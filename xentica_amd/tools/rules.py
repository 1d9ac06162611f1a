"""
Life-like rule encoding (``xentica/tools/rules.py:8-37``).

Birth counts occupy bits 0-8 and survival counts bits 9-17, so
``B3/S23`` is ``(1 << 3) | (1 << 11) | (1 << 12) = 6152``.
"""
import re

__all__ = ['LifeLike', ]


class LifeLike:
    """Conversions between Golly rule strings and integer rules."""

    @staticmethod
    def golly2int(golly_str):
        """``"B3/S23"`` -> ``6152``."""
        born, sus = re.sub("[^0-9/]", "", golly_str).split("/", 1)
        # a sum, not an OR: a repeated digit counts twice, as in the
        # reference
        born_rule = sum(1 << int(digit) for digit in born)
        sustained_rule = sum(1 << (int(digit) + 9) for digit in sus)
        return sustained_rule + born_rule

    @staticmethod
    def int2golly(rule):
        """``6152`` -> ``"B3/S23"``."""
        born = "".join(str(i) for i in range(9) if (rule >> i) & 1)
        sus = "".join(str(i) for i in range(9) if (rule >> (i + 9)) & 1)
        return "B%s/S%s" % (born, sus)
