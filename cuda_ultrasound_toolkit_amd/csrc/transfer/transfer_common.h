/*
 * Names shared by the transfer server and client: the command socket path
 * and the two shared-memory segment names derived from one base name
 * (CUT_TRANSFER_DEFAULT_NAME, the CUT_TRANSFER_NAME environment variable on
 * the client side, --name on the server side).
 */
#ifndef CUT_TRANSFER_COMMON_H
#define CUT_TRANSFER_COMMON_H

#include <stdio.h>
#include <string.h>

#include "cut_transfer.h"

#define CUT_TRANSFER_NAME_MAX 64

typedef struct CutTransferNames {
    char socket_path[108]; /* sizeof(sockaddr_un.sun_path) */
    char data[CUT_TRANSFER_NAME_MAX + 16];
    char params[CUT_TRANSFER_NAME_MAX + 16];
} CutTransferNames;

/* 0 when the base name is empty, too long or holds a '/'. */
static inline int cut_transfer_names(const char* base, CutTransferNames* n)
{
    const size_t len = base ? strlen(base) : 0;
    if (len == 0 || len > CUT_TRANSFER_NAME_MAX || strchr(base, '/')) return 0;
    snprintf(n->socket_path, sizeof(n->socket_path), "/tmp/%s_command.sock", base);
    snprintf(n->data, sizeof(n->data), "/%s_data", base);
    snprintf(n->params, sizeof(n->params), "/%s_parameters", base);
    return 1;
}

#endif /* CUT_TRANSFER_COMMON_H */
