/*
 * parlite.h -- command-line parameters and diagnostics for the drop-in programs.
 *
 * Same calling convention as the Seismic Unix getpar family the reference programs
 * use (initargs/getparint/getparfloat/getparstring/requestdoc/err/warn, e.g.
 * src/main/sufdmod2dTTIelastic/sufdmod2dTTIelastic_bond_transformation.c:118-294):
 * parameters are "key=value" words on the command line, optionally also read from
 * files named by "par=file" (one or more key=value per line, '#' starts a comment);
 * when a key is given more than once, the last occurrence wins.
 */
#ifndef PARLITE_H
#define PARLITE_H

void initargs(int argc, char **argv);
/* Print the self-documentation and exit when the program is run without parameters. */
void requestdoc(int flag);
void set_selfdoc(const char **sdoc);
int getparint(const char *name, int *v);
int getparfloat(const char *name, float *v);
int getparstring(const char *name, char **v);
const char *progname(void);

/* Print "<prog>: <message>" to stderr; err() then exits with status 1. */
void err(const char *fmt, ...) __attribute__((format(printf, 1, 2), noreturn));
void warn(const char *fmt, ...) __attribute__((format(printf, 1, 2)));

#endif
