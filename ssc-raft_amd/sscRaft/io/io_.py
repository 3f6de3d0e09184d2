"""Dictionary defaults, as sscRaft/io/io_.py:132-183 (SetDictionary)."""
import logging

logger = logging.getLogger("sscRaft")

__all__ = ["SetDictionary"]


def _required(dic, keys):
    for key in keys:
        if key not in dic:
            msg = f"Missing required dictionary key: {key}"
            logger.info(msg)
            raise ValueError(msg)


def _default(dic, key, value):
    if key not in dic:
        logger.info(f"Using default {key}: {value}.")
        dic[key] = value
    elif isinstance(dic[key], list) and isinstance(value, (list, tuple)):
        # list-valued keys: fill missing trailing entries from the default
        for i in range(len(dic[key]), len(value)):
            dic[key].append(value[i])


def SetDictionary(dic, required, optional, default):
    """Check `required` keys (ValueError if one is missing) and fill absent
    `optional` keys with `default`, in place; returns `dic`."""
    if dic is None:
        dic = {}
    _required(dic, required)
    for key, value in zip(optional, default):
        _default(dic, key, value)
    return dic
