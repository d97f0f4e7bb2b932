#!/usr/bin/env python3
"""Write minknow_data_fields.json: the wire interface of MinKNOW's Read Until
data service, as the reference's vendored minknow_api 5.4 defines it.

The reference ships the generated module
src/ReadUntil/read_until_env/lib/python3.6/site-packages/minknow_api/data_pb2.py,
whose serialized FileDescriptorProto is the data.proto schema.  This script
parses that module as text (ast; nothing of it is imported or run), decodes
the descriptor with protobuf, and records, for the messages and methods a
Read Until client uses, every field's number, name, type, label, type name
and oneof.  The fixture is data: tests compare dtwax_amd.minknow's own
protocol declaration against it.

Run where the reference tree is mounted at /root/reference.
"""
import ast
import json
import os
import sys

from google.protobuf import descriptor_pb2

REF = "/root/reference"
SITE = "src/ReadUntil/read_until_env/lib/python3.6/site-packages/minknow_api/"
PB2 = SITE + "data_pb2.py"
STATS = SITE + "statistics_pb2.py"  # the ReadEndReason enum that ReadData refers to
HERE = os.path.dirname(os.path.abspath(__file__))
MESSAGES = ("GetLiveReadsRequest", "GetLiveReadsResponse", "GetDataTypesRequest",
            "GetDataTypesResponse")
METHODS = ("get_live_reads", "get_data_types")


def _message(m):
    T = descriptor_pb2.FieldDescriptorProto
    return {
        "fields": [{"number": f.number, "name": f.name, "type": T.Type.Name(f.type),
                    "label": T.Label.Name(f.label), "type_name": f.type_name,
                    "oneof": (m.oneof_decl[f.oneof_index].name
                              if f.HasField("oneof_index") else None)}
                   for f in m.field],
        "nested": {n.name: _message(n) for n in m.nested_type},
        "enums": {e.name: {v.name: v.number for v in e.value} for e in m.enum_type},
        "map_entry": bool(m.options.map_entry),
    }


def _descriptor(rel):
    """The FileDescriptorProto literal of a generated module, parsed as text."""
    for node in ast.walk(ast.parse(open(os.path.join(REF, rel)).read())):
        if isinstance(node, ast.Call) and getattr(node.func, "attr", "") == "AddSerializedFile":
            return descriptor_pb2.FileDescriptorProto.FromString(node.args[0].value)
    raise SystemExit(f"no serialized descriptor in {rel}")


def main():
    if not os.path.isfile(os.path.join(REF, PB2)):
        sys.exit("reference tree not mounted; the fixture is already committed")
    fd = _descriptor(PB2)
    st = _descriptor(STATS)
    out = {"source": PB2, "package": fd.package, "syntax": fd.syntax,
           "read_end_reason": {v.name: v.number for e in st.enum_type
                               if e.name == "ReadEndReason" for v in e.value},
           "messages": {m.name: _message(m) for m in fd.message_type if m.name in MESSAGES},
           "methods": {}}
    for s in fd.service:
        for meth in s.method:
            if meth.name in METHODS:
                out["methods"][meth.name] = {
                    "service": f"{fd.package}.{s.name}", "input": meth.input_type,
                    "output": meth.output_type, "client_streaming": meth.client_streaming,
                    "server_streaming": meth.server_streaming}
    with open(os.path.join(HERE, "minknow_data_fields.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
