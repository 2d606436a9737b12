## Argument checks with the messages of sparseDist's own validators
## (R/utils-validate.R:17-42 in the reference), so scripts that catch them
## keep working.

.chk_flag <- function(x, name) {
  if (!is.logical(x) || length(x) != 1L || is.na(x))
    stop(sprintf("'%s' must be TRUE or FALSE.", name), call. = FALSE)
  invisible(x)
}

.chk_count <- function(x, name, allow_zero = FALSE) {
  lower <- if (allow_zero) 0 else 1
  ok <- is.numeric(x) && length(x) == 1L && !is.na(x) && is.finite(x) &&
    x == floor(x) && x >= lower && x <= .Machine$integer.max
  if (!ok)
    stop(sprintf("'%s' must be a %s integer no larger than .Machine$integer.max.", name,
                 if (allow_zero) "non-negative" else "positive"), call. = FALSE)
  as.integer(x)
}

.methods <- c("binary", "jaccard", "cosine", "euclidean", "manhattan", "pearson", "js",
              "covariance")

## method code of the C ABI (include/sd.h); "jaccard" is an alias of binary
.method_code <- function(method) {
  switch(method, binary = 0L, jaccard = 0L, cosine = 1L, euclidean = 2L, manhattan = 3L,
         pearson = 4L, covariance = 5L,
         stop("method \"js\" (dense Jensen-Shannon) is not provided on the GPU.", call. = FALSE))
}

.check_direction <- function(method, dist) {
  if (method %in% c("euclidean", "manhattan", "js") && !dist)
    stop(sprintf(paste0("'dist = FALSE' is only meaningful for methods \"binary\", \"cosine\" ",
                        "and \"pearson\"; %s has no similarity form."), method), call. = FALSE)
}

## dense matrices and other sparse classes become dgCMatrix, as the
## reference's .as_csparse does (R/sparseDist.R:198-202)
.as_dgc <- function(M) {
  if (is(M, "dgCMatrix")) return(M)
  if (is.data.frame(M)) M <- as.matrix(M)
  as(as(as(M, "CsparseMatrix"), "generalMatrix"), "dMatrix")
}
