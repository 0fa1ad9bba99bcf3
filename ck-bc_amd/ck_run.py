"""Runs a CK program entry the way `ck run program:<name>` does, without the
CK framework: expands the $#...#$ variables of run_cmd_main (bin file,
dataset path and file name, script extension), runs the compile script in
the program's tmp/ directory, executes the command there and honours
ignore_return_code.

    python ck_run.py program/bc-hip --dataset /path/to/facebook_combined.txt
"""
import argparse
import json
import os
import shlex
import subprocess
import sys


def expand(cmd, env):
    for k, v in env.items():
        cmd = cmd.replace("$#" + k + "#$", v)
    return cmd


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("program", help="program entry directory (holds .cm/meta.json)")
    ap.add_argument("--dataset", required=True, help="graph file selected for the run")
    ap.add_argument("--cmd", default="default")
    ap.add_argument("--extra", default="", help="extra command-line arguments")
    a = ap.parse_args(argv)
    prog = os.path.abspath(a.program)
    meta = json.load(open(os.path.join(prog, ".cm", "meta.json")))
    tmp = os.path.join(prog, "tmp")
    os.makedirs(tmp, exist_ok=True)
    env = dict(os.environ, CK_PROG_TARGET_EXE=meta["target_file"])
    if meta.get("use_compile_script") == "yes":
        cc = expand(meta["compile_cmds"]["default"]["cmd"], {"script_ext": ".sh"})
        subprocess.check_call(shlex.split(cc), cwd=tmp, env=env)
    rc_meta = meta["run_cmds"][a.cmd]
    d = os.path.abspath(a.dataset)
    cmd = expand(rc_meta["run_time"]["run_cmd_main"], {
        "BIN_FILE": "./" + meta["target_file"],
        "dataset_path": os.path.dirname(d) + os.sep,
        "dataset_filename": os.path.basename(d),
    })
    cmd = cmd + (" " + a.extra if a.extra else "")
    print("ck_run:", cmd, flush=True)
    rc = subprocess.call(shlex.split(cmd), cwd=tmp, env=env)
    if rc_meta.get("ignore_return_code", "no") == "no" and rc != 0:
        print(f"ck_run: program returned {rc}", file=sys.stderr)
        return rc
    return 0


if __name__ == "__main__":
    sys.exit(main())
