__all__ = ["xform", "abi"]
