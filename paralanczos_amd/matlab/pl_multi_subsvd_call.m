function [eigs, det1, err1] = pl_multi_subsvd_call(g, subIndex, minor)
%PL_MULTI_SUBSVD_CALL  One call of pl_multi_subsvd in libparalanczos.so:
%   all eigenvalues (ascending per subset), determinant and error flag of
%   the length(subIndex)/minor principal submatrices of g whose 0-based
%   indices subIndex lists, minor at a time. Host arrays in and out (the
%   reference moved gpuArrays through the Parallel Computing Toolbox).
%   Returns -1, -1, -1 if the library reports a failure.
pl_loadlibrary();
g = double(g);
idx = int32(subIndex(:));
L = numel(idx);
M = size(g, 1);
nsub = L / minor;
eigsP = libpointer('doublePtr', zeros(L, 1));
detP = libpointer('doublePtr', zeros(nsub, 1));
errP = libpointer('int32Ptr', zeros(nsub, 1, 'int32'));
rc = calllib('libparalanczos', 'pl_multi_subsvd', g, M, idx, L, minor, eigsP, detP, errP);
if rc ~= 0
    eigs = -1; det1 = -1; err1 = -1;
    return;
end
eigs = eigsP.Value;
det1 = detP.Value;
err1 = errP.Value;
end
