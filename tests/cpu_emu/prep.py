"""Rewrites a .hip source for the CPU emulation: every <<<grid, block, lds,
stream>>> launch becomes emu_launch(...) and the dynamic LDS array becomes
the emulator's per-block buffer.  Usage: prep.py in.hip out.cpp"""
import re, sys
src = open(sys.argv[1]).read()
src = src.replace('#include "../../include/lcodepy_amd.h"', '#include "lcodepy_amd.h"')
src = src.replace('extern __shared__ __align__(16) unsigned char lds_raw[];',
                  'unsigned char *lds_raw = emu_lds;')
def split_top(s):
    out, depth, cur = [], 0, ''
    for ch in s:
        if ch in '(<[': depth += 1
        if ch in ')>]': depth -= 1
        if ch == ',' and depth == 0:
            out.append(cur.strip()); cur = ''
        else:
            cur += ch
    out.append(cur.strip())
    return out
pat = re.compile(r'([A-Za-z_][A-Za-z_0-9:]*(?:<[^;(){}&]*?>)?)\s*<<<(.*?)>>>\((.*?)\);', re.S)
def rep(m):
    name, cfg, args = m.group(1), m.group(2), m.group(3)
    c = split_top(cfg)
    while len(c) < 4: c.append('0')
    return f'emu_launch(dim3({c[0]}), dim3({c[1]}), {c[2]}, [&] {{ {name}({args}); }});'
src = pat.sub(rep, src)
src = 'extern thread_local unsigned char *emu_lds;\n' + src if 'emu_lds' in src else src
open(sys.argv[2], 'w').write(src)
