"""Writes rd_jit_src.inc: the device headers as C++ raw string literals, so
the library can compile a dictionary-specialised kernel with hiprtc at run
time.  Usage: python3 embed_headers.py out.inc header...  (build step)."""
import os
import sys

out, headers = sys.argv[1], sys.argv[2:]
with open(out, "w") as f:
    f.write("// generated by embed_headers.py from " + " ".join(headers) + "\n")
    f.write("static const int kJitNumHeaders = %d;\n" % len(headers))
    f.write("static const char *const kJitHeaderNames[] = {%s};\n"
            % ", ".join('"%s"' % os.path.basename(h) for h in headers))
    f.write("static const char *const kJitHeaders[] = {\n")
    for h in headers:
        text = open(h).read()
        assert ")RDJIT\"" not in text
        f.write('R"RDJIT(' + text + ')RDJIT",\n')
    f.write("};\n")
