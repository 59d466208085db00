function [eigs, det1, err1] = gpu_svd(g)
%GPU_SVD  All eigenvalues of the symmetric matrix g (dyadic size <= 128) on
%   the GPU: the identity index set through the subset path. Same signature,
%   checks and outputs as Matlab/gpu_svd.m:13-57.
eigs = -1; det1 = -1; err1 = -1;
[M, N] = size(g);
if M ~= N
    disp('dimensions of input don''t agree! ');
    return;
end
if floor(log2(M)) - log2(M) ~= 0
    disp('dimensions are not dyadic!');
    return;
end
if M > 128
    disp('Matrices larger than 128 x 128 are not supported!');
    return;
end
[eigs, det1, err1] = pl_multi_subsvd_call(g, 0:M-1, M);
end
